// Generated by gen_panel_asm.py (same directory); see its docstring.
// A segment of chunks of the FP32 panel multiply for one wave: per chunk the
// barrier, the next chunk's entry loads and staging, and the entry loop.
#pragma once
// FP32 D, 16 rows per wave: accumulators v0..v63, component c of row slot q
// in v[16 c + q]; scratch v64..v97
#define GCOO_PANEL_SEGMENT_ASM_R16 \
    "s_mov_b32 s28, m0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[64:67], v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[68:71], v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[72:75], v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[76:79], v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[80:83], v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v96\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_R16 \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", \
    "v76", "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", \
    "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95", "v96", "v97", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
#define GCOO_PANEL_ACC_OPERANDS_R16(a) \
    "+{v[0:7]}"(a[0]), "+{v[8:15]}"(a[1]), "+{v[16:23]}"(a[2]), "+{v[24:31]}"(a[3]), \
    "+{v[32:39]}"(a[4]), "+{v[40:47]}"(a[5]), "+{v[48:55]}"(a[6]), "+{v[56:63]}"(a[7])
// FP32 D, 21 rows per wave: accumulators v0..v83, component c of row slot q
// in v[21 c + q]; scratch v84..v117
#define GCOO_PANEL_SEGMENT_ASM_R21 \
    "s_mov_b32 s28, m0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[84:87], v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[88:91], v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[92:95], v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[96:99], v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[100:103], v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[104:107], v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[108:111], v116\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_R21 \
    "v84", "v85", "v86", "v87", "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95", \
    "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103", "v104", "v105", "v106", "v107", \
    "v108", "v109", "v110", "v111", "v112", "v113", "v114", "v115", "v116", "v117", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
#define GCOO_PANEL_ACC_OPERANDS_R21(a) \
    "+{v[0:3]}"(a[0]), "+{v[4:7]}"(a[1]), "+{v[8:11]}"(a[2]), "+{v[12:15]}"(a[3]), \
    "+{v[16:19]}"(a[4]), "+{v[20:23]}"(a[5]), "+{v[24:27]}"(a[6]), "+{v[28:31]}"(a[7]), \
    "+{v[32:35]}"(a[8]), "+{v[36:39]}"(a[9]), "+{v[40:43]}"(a[10]), "+{v[44:47]}"(a[11]), \
    "+{v[48:51]}"(a[12]), "+{v[52:55]}"(a[13]), "+{v[56:59]}"(a[14]), "+{v[60:63]}"(a[15]), \
    "+{v[64:67]}"(a[16]), "+{v[68:71]}"(a[17]), "+{v[72:75]}"(a[18]), "+{v[76:79]}"(a[19]), \
    "+{v[80:83]}"(a[20])
// FP32 D, 28 rows per wave: accumulators v0..v111, component c of row slot q
// in v[28 c + q]; scratch v112..v145
#define GCOO_PANEL_SEGMENT_ASM_R28 \
    "s_mov_b32 s28, m0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[140:143], v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[112:115], v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[116:119], v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[120:123], v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[124:127], v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[128:131], v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[132:135], v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[136:139], v144\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_R28 \
    "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", "v122", "v123", \
    "v124", "v125", "v126", "v127", "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135", \
    "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143", "v144", "v145", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
#define GCOO_PANEL_ACC_OPERANDS_R28(a) \
    "+{v[0:7]}"(a[0]), "+{v[8:15]}"(a[1]), "+{v[16:23]}"(a[2]), "+{v[24:31]}"(a[3]), \
    "+{v[32:39]}"(a[4]), "+{v[40:47]}"(a[5]), "+{v[48:55]}"(a[6]), "+{v[56:63]}"(a[7]), \
    "+{v[64:71]}"(a[8]), "+{v[72:79]}"(a[9]), "+{v[80:87]}"(a[10]), "+{v[88:95]}"(a[11]), \
    "+{v[96:103]}"(a[12]), "+{v[104:111]}"(a[13])
// FP32 D, 42 rows per wave: accumulators v0..v167, component c of row slot q
// in v[42 c + q]; scratch v168..v201
#define GCOO_PANEL_SEGMENT_ASM_R42 \
    "s_mov_b32 s28, m0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 32\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v200, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v201, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v200, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v201, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v200, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v200, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v200, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v201, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v200, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v201, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v196, v0\n\t" \
    "v_fma_f32 v42, s51, v197, v42\n\t" \
    "v_fma_f32 v84, s51, v198, v84\n\t" \
    "v_fma_f32 v126, s51, v199, v126\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v200, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v196, v0\n\t" \
    "v_fma_f32 v42, s51, v197, v42\n\t" \
    "v_fma_f32 v84, s51, v198, v84\n\t" \
    "v_fma_f32 v126, s51, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v200, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v201, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v200, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v201, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v196, v0\n\t" \
    "v_fma_f32 v42, s51, v197, v42\n\t" \
    "v_fma_f32 v84, s51, v198, v84\n\t" \
    "v_fma_f32 v126, s51, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v196, v0\n\t" \
    "v_fma_f32 v42, s51, v197, v42\n\t" \
    "v_fma_f32 v84, s51, v198, v84\n\t" \
    "v_fma_f32 v126, s51, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v200, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v200, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v201, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v200, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v201, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v184, v0\n\t" \
    "v_fma_f32 v42, s61, v185, v42\n\t" \
    "v_fma_f32 v84, s61, v186, v84\n\t" \
    "v_fma_f32 v126, s61, v187, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v188, v0\n\t" \
    "v_fma_f32 v42, s63, v189, v42\n\t" \
    "v_fma_f32 v84, s63, v190, v84\n\t" \
    "v_fma_f32 v126, s63, v191, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v192, v0\n\t" \
    "v_fma_f32 v42, s65, v193, v42\n\t" \
    "v_fma_f32 v84, s65, v194, v84\n\t" \
    "v_fma_f32 v126, s65, v195, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v196, v0\n\t" \
    "v_fma_f32 v42, s67, v197, v42\n\t" \
    "v_fma_f32 v84, s67, v198, v84\n\t" \
    "v_fma_f32 v126, s67, v199, v126\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v184, v0\n\t" \
    "v_fma_f32 v42, s61, v185, v42\n\t" \
    "v_fma_f32 v84, s61, v186, v84\n\t" \
    "v_fma_f32 v126, s61, v187, v126\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v188, v0\n\t" \
    "v_fma_f32 v42, s63, v189, v42\n\t" \
    "v_fma_f32 v84, s63, v190, v84\n\t" \
    "v_fma_f32 v126, s63, v191, v126\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v192, v0\n\t" \
    "v_fma_f32 v42, s65, v193, v42\n\t" \
    "v_fma_f32 v84, s65, v194, v84\n\t" \
    "v_fma_f32 v126, s65, v195, v126\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v196, v0\n\t" \
    "v_fma_f32 v42, s67, v197, v42\n\t" \
    "v_fma_f32 v84, s67, v198, v84\n\t" \
    "v_fma_f32 v126, s67, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 42\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v200, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "v_mad_u32_u16 v200, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v196, v0\n\t" \
    "v_fma_f32 v42, s51, v197, v42\n\t" \
    "v_fma_f32 v84, s51, v198, v84\n\t" \
    "v_fma_f32 v126, s51, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 42\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v200, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "v_mad_u32_u16 v200, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v184, v0\n\t" \
    "v_fma_f32 v42, s61, v185, v42\n\t" \
    "v_fma_f32 v84, s61, v186, v84\n\t" \
    "v_fma_f32 v126, s61, v187, v126\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v188, v0\n\t" \
    "v_fma_f32 v42, s63, v189, v42\n\t" \
    "v_fma_f32 v84, s63, v190, v84\n\t" \
    "v_fma_f32 v126, s63, v191, v126\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v192, v0\n\t" \
    "v_fma_f32 v42, s65, v193, v42\n\t" \
    "v_fma_f32 v84, s65, v194, v84\n\t" \
    "v_fma_f32 v126, s65, v195, v126\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v196, v0\n\t" \
    "v_fma_f32 v42, s67, v197, v42\n\t" \
    "v_fma_f32 v84, s67, v198, v84\n\t" \
    "v_fma_f32 v126, s67, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v200, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v201, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v200, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v201, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v200, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v201, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v200, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v168, v0\n\t" \
    "v_fma_f32 v42, s37, v169, v42\n\t" \
    "v_fma_f32 v84, s37, v170, v84\n\t" \
    "v_fma_f32 v126, s37, v171, v126\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v172, v0\n\t" \
    "v_fma_f32 v42, s39, v173, v42\n\t" \
    "v_fma_f32 v84, s39, v174, v84\n\t" \
    "v_fma_f32 v126, s39, v175, v126\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v176, v0\n\t" \
    "v_fma_f32 v42, s41, v177, v42\n\t" \
    "v_fma_f32 v84, s41, v178, v84\n\t" \
    "v_fma_f32 v126, s41, v179, v126\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v180, v0\n\t" \
    "v_fma_f32 v42, s43, v181, v42\n\t" \
    "v_fma_f32 v84, s43, v182, v84\n\t" \
    "v_fma_f32 v126, s43, v183, v126\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v184, v0\n\t" \
    "v_fma_f32 v42, s45, v185, v42\n\t" \
    "v_fma_f32 v84, s45, v186, v84\n\t" \
    "v_fma_f32 v126, s45, v187, v126\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v188, v0\n\t" \
    "v_fma_f32 v42, s47, v189, v42\n\t" \
    "v_fma_f32 v84, s47, v190, v84\n\t" \
    "v_fma_f32 v126, s47, v191, v126\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v192, v0\n\t" \
    "v_fma_f32 v42, s49, v193, v42\n\t" \
    "v_fma_f32 v84, s49, v194, v84\n\t" \
    "v_fma_f32 v126, s49, v195, v126\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v200, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v201, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v200, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v201, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v200, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v201, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v200, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v168, v0\n\t" \
    "v_fma_f32 v42, s53, v169, v42\n\t" \
    "v_fma_f32 v84, s53, v170, v84\n\t" \
    "v_fma_f32 v126, s53, v171, v126\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v172, v0\n\t" \
    "v_fma_f32 v42, s55, v173, v42\n\t" \
    "v_fma_f32 v84, s55, v174, v84\n\t" \
    "v_fma_f32 v126, s55, v175, v126\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v176, v0\n\t" \
    "v_fma_f32 v42, s57, v177, v42\n\t" \
    "v_fma_f32 v84, s57, v178, v84\n\t" \
    "v_fma_f32 v126, s57, v179, v126\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v180, v0\n\t" \
    "v_fma_f32 v42, s59, v181, v42\n\t" \
    "v_fma_f32 v84, s59, v182, v84\n\t" \
    "v_fma_f32 v126, s59, v183, v126\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v184, v0\n\t" \
    "v_fma_f32 v42, s61, v185, v42\n\t" \
    "v_fma_f32 v84, s61, v186, v84\n\t" \
    "v_fma_f32 v126, s61, v187, v126\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v188, v0\n\t" \
    "v_fma_f32 v42, s63, v189, v42\n\t" \
    "v_fma_f32 v84, s63, v190, v84\n\t" \
    "v_fma_f32 v126, s63, v191, v126\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v192, v0\n\t" \
    "v_fma_f32 v42, s65, v193, v42\n\t" \
    "v_fma_f32 v84, s65, v194, v84\n\t" \
    "v_fma_f32 v126, s65, v195, v126\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 32\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v200, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v201, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v200, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v201, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v200, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v200, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v200, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v201, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v200, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v201, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v196, v0\n\t" \
    "v_fma_f32 v42, s83, v197, v42\n\t" \
    "v_fma_f32 v84, s83, v198, v84\n\t" \
    "v_fma_f32 v126, s83, v199, v126\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v200, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v196, v0\n\t" \
    "v_fma_f32 v42, s83, v197, v42\n\t" \
    "v_fma_f32 v84, s83, v198, v84\n\t" \
    "v_fma_f32 v126, s83, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v200, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v201, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v200, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v201, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v196, v0\n\t" \
    "v_fma_f32 v42, s83, v197, v42\n\t" \
    "v_fma_f32 v84, s83, v198, v84\n\t" \
    "v_fma_f32 v126, s83, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v196, v0\n\t" \
    "v_fma_f32 v42, s83, v197, v42\n\t" \
    "v_fma_f32 v84, s83, v198, v84\n\t" \
    "v_fma_f32 v126, s83, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v200, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(4)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v200, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v201, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v200, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v201, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v184, v0\n\t" \
    "v_fma_f32 v42, s93, v185, v42\n\t" \
    "v_fma_f32 v84, s93, v186, v84\n\t" \
    "v_fma_f32 v126, s93, v187, v126\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v188, v0\n\t" \
    "v_fma_f32 v42, s95, v189, v42\n\t" \
    "v_fma_f32 v84, s95, v190, v84\n\t" \
    "v_fma_f32 v126, s95, v191, v126\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v192, v0\n\t" \
    "v_fma_f32 v42, s97, v193, v42\n\t" \
    "v_fma_f32 v84, s97, v194, v84\n\t" \
    "v_fma_f32 v126, s97, v195, v126\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v196, v0\n\t" \
    "v_fma_f32 v42, s99, v197, v42\n\t" \
    "v_fma_f32 v84, s99, v198, v84\n\t" \
    "v_fma_f32 v126, s99, v199, v126\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v184, v0\n\t" \
    "v_fma_f32 v42, s93, v185, v42\n\t" \
    "v_fma_f32 v84, s93, v186, v84\n\t" \
    "v_fma_f32 v126, s93, v187, v126\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v188, v0\n\t" \
    "v_fma_f32 v42, s95, v189, v42\n\t" \
    "v_fma_f32 v84, s95, v190, v84\n\t" \
    "v_fma_f32 v126, s95, v191, v126\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v192, v0\n\t" \
    "v_fma_f32 v42, s97, v193, v42\n\t" \
    "v_fma_f32 v84, s97, v194, v84\n\t" \
    "v_fma_f32 v126, s97, v195, v126\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v196, v0\n\t" \
    "v_fma_f32 v42, s99, v197, v42\n\t" \
    "v_fma_f32 v84, s99, v198, v84\n\t" \
    "v_fma_f32 v126, s99, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 42\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v200, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "v_mad_u32_u16 v200, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v196, v0\n\t" \
    "v_fma_f32 v42, s83, v197, v42\n\t" \
    "v_fma_f32 v84, s83, v198, v84\n\t" \
    "v_fma_f32 v126, s83, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 42\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v200, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "v_mad_u32_u16 v201, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "v_mad_u32_u16 v200, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "v_mad_u32_u16 v201, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "v_mad_u32_u16 v200, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "v_mad_u32_u16 v201, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "v_mad_u32_u16 v200, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    "v_mad_u32_u16 v201, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[196:199], v201\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v184, v0\n\t" \
    "v_fma_f32 v42, s93, v185, v42\n\t" \
    "v_fma_f32 v84, s93, v186, v84\n\t" \
    "v_fma_f32 v126, s93, v187, v126\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v188, v0\n\t" \
    "v_fma_f32 v42, s95, v189, v42\n\t" \
    "v_fma_f32 v84, s95, v190, v84\n\t" \
    "v_fma_f32 v126, s95, v191, v126\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v192, v0\n\t" \
    "v_fma_f32 v42, s97, v193, v42\n\t" \
    "v_fma_f32 v84, s97, v194, v84\n\t" \
    "v_fma_f32 v126, s97, v195, v126\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v196, v0\n\t" \
    "v_fma_f32 v42, s99, v197, v42\n\t" \
    "v_fma_f32 v84, s99, v198, v84\n\t" \
    "v_fma_f32 v126, s99, v199, v126\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v200, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v201, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v200, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v201, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v200, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v201, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v200, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v168, v0\n\t" \
    "v_fma_f32 v42, s69, v169, v42\n\t" \
    "v_fma_f32 v84, s69, v170, v84\n\t" \
    "v_fma_f32 v126, s69, v171, v126\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v172, v0\n\t" \
    "v_fma_f32 v42, s71, v173, v42\n\t" \
    "v_fma_f32 v84, s71, v174, v84\n\t" \
    "v_fma_f32 v126, s71, v175, v126\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v176, v0\n\t" \
    "v_fma_f32 v42, s73, v177, v42\n\t" \
    "v_fma_f32 v84, s73, v178, v84\n\t" \
    "v_fma_f32 v126, s73, v179, v126\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v180, v0\n\t" \
    "v_fma_f32 v42, s75, v181, v42\n\t" \
    "v_fma_f32 v84, s75, v182, v84\n\t" \
    "v_fma_f32 v126, s75, v183, v126\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v184, v0\n\t" \
    "v_fma_f32 v42, s77, v185, v42\n\t" \
    "v_fma_f32 v84, s77, v186, v84\n\t" \
    "v_fma_f32 v126, s77, v187, v126\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v188, v0\n\t" \
    "v_fma_f32 v42, s79, v189, v42\n\t" \
    "v_fma_f32 v84, s79, v190, v84\n\t" \
    "v_fma_f32 v126, s79, v191, v126\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v192, v0\n\t" \
    "v_fma_f32 v42, s81, v193, v42\n\t" \
    "v_fma_f32 v84, s81, v194, v84\n\t" \
    "v_fma_f32 v126, s81, v195, v126\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v200, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[168:171], v200\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v201, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[172:175], v201\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v200, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[176:179], v200\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v201, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[180:183], v201\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v200, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[184:187], v200\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v201, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[188:191], v201\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v200, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b128 v[192:195], v200\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v168, v0\n\t" \
    "v_fma_f32 v42, s85, v169, v42\n\t" \
    "v_fma_f32 v84, s85, v170, v84\n\t" \
    "v_fma_f32 v126, s85, v171, v126\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v172, v0\n\t" \
    "v_fma_f32 v42, s87, v173, v42\n\t" \
    "v_fma_f32 v84, s87, v174, v84\n\t" \
    "v_fma_f32 v126, s87, v175, v126\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v176, v0\n\t" \
    "v_fma_f32 v42, s89, v177, v42\n\t" \
    "v_fma_f32 v84, s89, v178, v84\n\t" \
    "v_fma_f32 v126, s89, v179, v126\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v180, v0\n\t" \
    "v_fma_f32 v42, s91, v181, v42\n\t" \
    "v_fma_f32 v84, s91, v182, v84\n\t" \
    "v_fma_f32 v126, s91, v183, v126\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v184, v0\n\t" \
    "v_fma_f32 v42, s93, v185, v42\n\t" \
    "v_fma_f32 v84, s93, v186, v84\n\t" \
    "v_fma_f32 v126, s93, v187, v126\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v188, v0\n\t" \
    "v_fma_f32 v42, s95, v189, v42\n\t" \
    "v_fma_f32 v84, s95, v190, v84\n\t" \
    "v_fma_f32 v126, s95, v191, v126\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v192, v0\n\t" \
    "v_fma_f32 v42, s97, v193, v42\n\t" \
    "v_fma_f32 v84, s97, v194, v84\n\t" \
    "v_fma_f32 v126, s97, v195, v126\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_R42 \
    "v168", "v169", "v170", "v171", "v172", "v173", "v174", "v175", "v176", "v177", "v178", "v179", \
    "v180", "v181", "v182", "v183", "v184", "v185", "v186", "v187", "v188", "v189", "v190", "v191", \
    "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199", "v200", "v201", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
#define GCOO_PANEL_ACC_OPERANDS_R42(a) \
    "+{v[0:7]}"(a[0]), "+{v[8:15]}"(a[1]), "+{v[16:23]}"(a[2]), "+{v[24:31]}"(a[3]), \
    "+{v[32:39]}"(a[4]), "+{v[40:47]}"(a[5]), "+{v[48:55]}"(a[6]), "+{v[56:63]}"(a[7]), \
    "+{v[64:71]}"(a[8]), "+{v[72:79]}"(a[9]), "+{v[80:87]}"(a[10]), "+{v[88:95]}"(a[11]), \
    "+{v[96:103]}"(a[12]), "+{v[104:111]}"(a[13]), "+{v[112:119]}"(a[14]), "+{v[120:127]}"(a[15]), \
    "+{v[128:135]}"(a[16]), "+{v[136:143]}"(a[17]), "+{v[144:151]}"(a[18]), "+{v[152:159]}"(a[19]), \
    "+{v[160:167]}"(a[20])
// BF16 D, 16 rows per wave: accumulators v0..v63, component c of row slot q
// in v[16 c + q]; scratch v64..v97
#define GCOO_PANEL_SEGMENT_ASM_BF16_R16 \
    "s_mov_b32 s28, m0\n\t" \
    "v_mov_b32 v64, 0\n\t" \
    "v_mov_b32 v68, 0\n\t" \
    "v_mov_b32 v72, 0\n\t" \
    "v_mov_b32 v76, 0\n\t" \
    "v_mov_b32 v80, 0\n\t" \
    "v_mov_b32 v84, 0\n\t" \
    "v_mov_b32 v88, 0\n\t" \
    "v_mov_b32 v92, 0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz00_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn00_%=\n\t" \
    ".Lsz00_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn00_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz10_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn10_%=\n\t" \
    ".Lsz10_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn10_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz20_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn20_%=\n\t" \
    ".Lsz20_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn20_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz30_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn30_%=\n\t" \
    ".Lsz30_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn30_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz40_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn40_%=\n\t" \
    ".Lsz40_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn40_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz50_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn50_%=\n\t" \
    ".Lsz50_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn50_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz60_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn60_%=\n\t" \
    ".Lsz60_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn60_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz70_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn70_%=\n\t" \
    ".Lsz70_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn70_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz80_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn80_%=\n\t" \
    ".Lsz80_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn80_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz90_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn90_%=\n\t" \
    ".Lsz90_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn90_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz100_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn100_%=\n\t" \
    ".Lsz100_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn100_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz110_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn110_%=\n\t" \
    ".Lsz110_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn110_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz120_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn120_%=\n\t" \
    ".Lsz120_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn120_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz130_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn130_%=\n\t" \
    ".Lsz130_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn130_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz140_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn140_%=\n\t" \
    ".Lsz140_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn140_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    ".Lgvp0s0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    ".Lgvp1s0_%=:\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    ".Lgvp2s0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    ".Lgvp3s0_%=:\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v92, v0\n\t" \
    "v_fma_f32 v16, s51, v93, v16\n\t" \
    "v_fma_f32 v32, s51, v94, v32\n\t" \
    "v_fma_f32 v48, s51, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v92, v0\n\t" \
    "v_fma_f32 v16, s67, v93, v16\n\t" \
    "v_fma_f32 v32, s67, v94, v32\n\t" \
    "v_fma_f32 v48, s67, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v96, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v97, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v96, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    ".Lgva0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v64, v0\n\t" \
    "v_fma_f32 v16, s37, v65, v16\n\t" \
    "v_fma_f32 v32, s37, v66, v32\n\t" \
    "v_fma_f32 v48, s37, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v68, v0\n\t" \
    "v_fma_f32 v16, s39, v69, v16\n\t" \
    "v_fma_f32 v32, s39, v70, v32\n\t" \
    "v_fma_f32 v48, s39, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v72, v0\n\t" \
    "v_fma_f32 v16, s41, v73, v16\n\t" \
    "v_fma_f32 v32, s41, v74, v32\n\t" \
    "v_fma_f32 v48, s41, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v76, v0\n\t" \
    "v_fma_f32 v16, s43, v77, v16\n\t" \
    "v_fma_f32 v32, s43, v78, v32\n\t" \
    "v_fma_f32 v48, s43, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v80, v0\n\t" \
    "v_fma_f32 v16, s45, v81, v16\n\t" \
    "v_fma_f32 v32, s45, v82, v32\n\t" \
    "v_fma_f32 v48, s45, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v84, v0\n\t" \
    "v_fma_f32 v16, s47, v85, v16\n\t" \
    "v_fma_f32 v32, s47, v86, v32\n\t" \
    "v_fma_f32 v48, s47, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v88, v0\n\t" \
    "v_fma_f32 v16, s49, v89, v16\n\t" \
    "v_fma_f32 v32, s49, v90, v32\n\t" \
    "v_fma_f32 v48, s49, v91, v48\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v96, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v97, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v96, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    ".Lgvb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v64, v0\n\t" \
    "v_fma_f32 v16, s53, v65, v16\n\t" \
    "v_fma_f32 v32, s53, v66, v32\n\t" \
    "v_fma_f32 v48, s53, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v68, v0\n\t" \
    "v_fma_f32 v16, s55, v69, v16\n\t" \
    "v_fma_f32 v32, s55, v70, v32\n\t" \
    "v_fma_f32 v48, s55, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v72, v0\n\t" \
    "v_fma_f32 v16, s57, v73, v16\n\t" \
    "v_fma_f32 v32, s57, v74, v32\n\t" \
    "v_fma_f32 v48, s57, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v76, v0\n\t" \
    "v_fma_f32 v16, s59, v77, v16\n\t" \
    "v_fma_f32 v32, s59, v78, v32\n\t" \
    "v_fma_f32 v48, s59, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v80, v0\n\t" \
    "v_fma_f32 v16, s61, v81, v16\n\t" \
    "v_fma_f32 v32, s61, v82, v32\n\t" \
    "v_fma_f32 v48, s61, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v84, v0\n\t" \
    "v_fma_f32 v16, s63, v85, v16\n\t" \
    "v_fma_f32 v32, s63, v86, v32\n\t" \
    "v_fma_f32 v48, s63, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v88, v0\n\t" \
    "v_fma_f32 v16, s65, v89, v16\n\t" \
    "v_fma_f32 v32, s65, v90, v32\n\t" \
    "v_fma_f32 v48, s65, v91, v48\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz01_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn01_%=\n\t" \
    ".Lsz01_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn01_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz11_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn11_%=\n\t" \
    ".Lsz11_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn11_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz21_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn21_%=\n\t" \
    ".Lsz21_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn21_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz31_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn31_%=\n\t" \
    ".Lsz31_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn31_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz41_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn41_%=\n\t" \
    ".Lsz41_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn41_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz51_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn51_%=\n\t" \
    ".Lsz51_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn51_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz61_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn61_%=\n\t" \
    ".Lsz61_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn61_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz71_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn71_%=\n\t" \
    ".Lsz71_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn71_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz81_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn81_%=\n\t" \
    ".Lsz81_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn81_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz91_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn91_%=\n\t" \
    ".Lsz91_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn91_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz101_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn101_%=\n\t" \
    ".Lsz101_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn101_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz111_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn111_%=\n\t" \
    ".Lsz111_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn111_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz121_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn121_%=\n\t" \
    ".Lsz121_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn121_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz131_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn131_%=\n\t" \
    ".Lsz131_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn131_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz141_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn141_%=\n\t" \
    ".Lsz141_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn141_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    ".Lgvp0s1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    ".Lgvp1s1_%=:\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    ".Lgvp2s1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    ".Lgvp3s1_%=:\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v92, v0\n\t" \
    "v_fma_f32 v16, s83, v93, v16\n\t" \
    "v_fma_f32 v32, s83, v94, v32\n\t" \
    "v_fma_f32 v48, s83, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    "v_mad_u32_u16 v97, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v97\n\t" \
    "ds_read_u16_d16_hi v92, v97\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v92, v0\n\t" \
    "v_fma_f32 v16, s99, v93, v16\n\t" \
    "v_fma_f32 v32, s99, v94, v32\n\t" \
    "v_fma_f32 v48, s99, v95, v48\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v96, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v97, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v96, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    ".Lgva1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v64, v0\n\t" \
    "v_fma_f32 v16, s69, v65, v16\n\t" \
    "v_fma_f32 v32, s69, v66, v32\n\t" \
    "v_fma_f32 v48, s69, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v68, v0\n\t" \
    "v_fma_f32 v16, s71, v69, v16\n\t" \
    "v_fma_f32 v32, s71, v70, v32\n\t" \
    "v_fma_f32 v48, s71, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v72, v0\n\t" \
    "v_fma_f32 v16, s73, v73, v16\n\t" \
    "v_fma_f32 v32, s73, v74, v32\n\t" \
    "v_fma_f32 v48, s73, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v76, v0\n\t" \
    "v_fma_f32 v16, s75, v77, v16\n\t" \
    "v_fma_f32 v32, s75, v78, v32\n\t" \
    "v_fma_f32 v48, s75, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v80, v0\n\t" \
    "v_fma_f32 v16, s77, v81, v16\n\t" \
    "v_fma_f32 v32, s77, v82, v32\n\t" \
    "v_fma_f32 v48, s77, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v84, v0\n\t" \
    "v_fma_f32 v16, s79, v85, v16\n\t" \
    "v_fma_f32 v32, s79, v86, v32\n\t" \
    "v_fma_f32 v48, s79, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v88, v0\n\t" \
    "v_fma_f32 v16, s81, v89, v16\n\t" \
    "v_fma_f32 v32, s81, v90, v32\n\t" \
    "v_fma_f32 v48, s81, v91, v48\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v96, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[66:67], v96\n\t" \
    "ds_read_u16_d16_hi v64, v96\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[70:71], v97\n\t" \
    "ds_read_u16_d16_hi v68, v97\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[74:75], v96\n\t" \
    "ds_read_u16_d16_hi v72, v96\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[78:79], v97\n\t" \
    "ds_read_u16_d16_hi v76, v97\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[82:83], v96\n\t" \
    "ds_read_u16_d16_hi v80, v96\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v97, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v97\n\t" \
    "ds_read_u16_d16_hi v84, v97\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v96, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v96\n\t" \
    "ds_read_u16_d16_hi v88, v96\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v65, 0xffff0000, v66\n\t" \
    "v_lshlrev_b32_e32 v66, 16, v67\n\t" \
    "v_and_b32_e32 v67, 0xffff0000, v67\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v69, 0xffff0000, v70\n\t" \
    "v_lshlrev_b32_e32 v70, 16, v71\n\t" \
    "v_and_b32_e32 v71, 0xffff0000, v71\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v73, 0xffff0000, v74\n\t" \
    "v_lshlrev_b32_e32 v74, 16, v75\n\t" \
    "v_and_b32_e32 v75, 0xffff0000, v75\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v77, 0xffff0000, v78\n\t" \
    "v_lshlrev_b32_e32 v78, 16, v79\n\t" \
    "v_and_b32_e32 v79, 0xffff0000, v79\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v81, 0xffff0000, v82\n\t" \
    "v_lshlrev_b32_e32 v82, 16, v83\n\t" \
    "v_and_b32_e32 v83, 0xffff0000, v83\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    ".Lgvb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v64, v0\n\t" \
    "v_fma_f32 v16, s85, v65, v16\n\t" \
    "v_fma_f32 v32, s85, v66, v32\n\t" \
    "v_fma_f32 v48, s85, v67, v48\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v68, v0\n\t" \
    "v_fma_f32 v16, s87, v69, v16\n\t" \
    "v_fma_f32 v32, s87, v70, v32\n\t" \
    "v_fma_f32 v48, s87, v71, v48\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v72, v0\n\t" \
    "v_fma_f32 v16, s89, v73, v16\n\t" \
    "v_fma_f32 v32, s89, v74, v32\n\t" \
    "v_fma_f32 v48, s89, v75, v48\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v76, v0\n\t" \
    "v_fma_f32 v16, s91, v77, v16\n\t" \
    "v_fma_f32 v32, s91, v78, v32\n\t" \
    "v_fma_f32 v48, s91, v79, v48\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v80, v0\n\t" \
    "v_fma_f32 v16, s93, v81, v16\n\t" \
    "v_fma_f32 v32, s93, v82, v32\n\t" \
    "v_fma_f32 v48, s93, v83, v48\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v84, v0\n\t" \
    "v_fma_f32 v16, s95, v85, v16\n\t" \
    "v_fma_f32 v32, s95, v86, v32\n\t" \
    "v_fma_f32 v48, s95, v87, v48\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v88, v0\n\t" \
    "v_fma_f32 v16, s97, v89, v16\n\t" \
    "v_fma_f32 v32, s97, v90, v32\n\t" \
    "v_fma_f32 v48, s97, v91, v48\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_BF16_R16 \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", \
    "v76", "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", \
    "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95", "v96", "v97", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
// BF16 D, 21 rows per wave: accumulators v0..v83, component c of row slot q
// in v[21 c + q]; scratch v84..v117
#define GCOO_PANEL_SEGMENT_ASM_BF16_R21 \
    "s_mov_b32 s28, m0\n\t" \
    "v_mov_b32 v84, 0\n\t" \
    "v_mov_b32 v88, 0\n\t" \
    "v_mov_b32 v92, 0\n\t" \
    "v_mov_b32 v96, 0\n\t" \
    "v_mov_b32 v100, 0\n\t" \
    "v_mov_b32 v104, 0\n\t" \
    "v_mov_b32 v108, 0\n\t" \
    "v_mov_b32 v112, 0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz00_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn00_%=\n\t" \
    ".Lsz00_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn00_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz10_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn10_%=\n\t" \
    ".Lsz10_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn10_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz20_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn20_%=\n\t" \
    ".Lsz20_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn20_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz30_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn30_%=\n\t" \
    ".Lsz30_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn30_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz40_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn40_%=\n\t" \
    ".Lsz40_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn40_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz50_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn50_%=\n\t" \
    ".Lsz50_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn50_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz60_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn60_%=\n\t" \
    ".Lsz60_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn60_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz70_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn70_%=\n\t" \
    ".Lsz70_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn70_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz80_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn80_%=\n\t" \
    ".Lsz80_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn80_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz90_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn90_%=\n\t" \
    ".Lsz90_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn90_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz100_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn100_%=\n\t" \
    ".Lsz100_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn100_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz110_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn110_%=\n\t" \
    ".Lsz110_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn110_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz120_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn120_%=\n\t" \
    ".Lsz120_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn120_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz130_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn130_%=\n\t" \
    ".Lsz130_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn130_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz140_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn140_%=\n\t" \
    ".Lsz140_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn140_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    ".Lgvp0s0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    ".Lgvp1s0_%=:\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    ".Lgvp2s0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    ".Lgvp3s0_%=:\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v112, v0\n\t" \
    "v_fma_f32 v21, s51, v113, v21\n\t" \
    "v_fma_f32 v42, s51, v114, v42\n\t" \
    "v_fma_f32 v63, s51, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v112, v0\n\t" \
    "v_fma_f32 v21, s67, v113, v21\n\t" \
    "v_fma_f32 v42, s67, v114, v42\n\t" \
    "v_fma_f32 v63, s67, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v116, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v117, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v116, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    ".Lgva0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v84, v0\n\t" \
    "v_fma_f32 v21, s37, v85, v21\n\t" \
    "v_fma_f32 v42, s37, v86, v42\n\t" \
    "v_fma_f32 v63, s37, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v88, v0\n\t" \
    "v_fma_f32 v21, s39, v89, v21\n\t" \
    "v_fma_f32 v42, s39, v90, v42\n\t" \
    "v_fma_f32 v63, s39, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v92, v0\n\t" \
    "v_fma_f32 v21, s41, v93, v21\n\t" \
    "v_fma_f32 v42, s41, v94, v42\n\t" \
    "v_fma_f32 v63, s41, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v96, v0\n\t" \
    "v_fma_f32 v21, s43, v97, v21\n\t" \
    "v_fma_f32 v42, s43, v98, v42\n\t" \
    "v_fma_f32 v63, s43, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v100, v0\n\t" \
    "v_fma_f32 v21, s45, v101, v21\n\t" \
    "v_fma_f32 v42, s45, v102, v42\n\t" \
    "v_fma_f32 v63, s45, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v104, v0\n\t" \
    "v_fma_f32 v21, s47, v105, v21\n\t" \
    "v_fma_f32 v42, s47, v106, v42\n\t" \
    "v_fma_f32 v63, s47, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v108, v0\n\t" \
    "v_fma_f32 v21, s49, v109, v21\n\t" \
    "v_fma_f32 v42, s49, v110, v42\n\t" \
    "v_fma_f32 v63, s49, v111, v63\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v116, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v117, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v116, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    ".Lgvb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v84, v0\n\t" \
    "v_fma_f32 v21, s53, v85, v21\n\t" \
    "v_fma_f32 v42, s53, v86, v42\n\t" \
    "v_fma_f32 v63, s53, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v88, v0\n\t" \
    "v_fma_f32 v21, s55, v89, v21\n\t" \
    "v_fma_f32 v42, s55, v90, v42\n\t" \
    "v_fma_f32 v63, s55, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v92, v0\n\t" \
    "v_fma_f32 v21, s57, v93, v21\n\t" \
    "v_fma_f32 v42, s57, v94, v42\n\t" \
    "v_fma_f32 v63, s57, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v96, v0\n\t" \
    "v_fma_f32 v21, s59, v97, v21\n\t" \
    "v_fma_f32 v42, s59, v98, v42\n\t" \
    "v_fma_f32 v63, s59, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v100, v0\n\t" \
    "v_fma_f32 v21, s61, v101, v21\n\t" \
    "v_fma_f32 v42, s61, v102, v42\n\t" \
    "v_fma_f32 v63, s61, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v104, v0\n\t" \
    "v_fma_f32 v21, s63, v105, v21\n\t" \
    "v_fma_f32 v42, s63, v106, v42\n\t" \
    "v_fma_f32 v63, s63, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v108, v0\n\t" \
    "v_fma_f32 v21, s65, v109, v21\n\t" \
    "v_fma_f32 v42, s65, v110, v42\n\t" \
    "v_fma_f32 v63, s65, v111, v63\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz01_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn01_%=\n\t" \
    ".Lsz01_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn01_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz11_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn11_%=\n\t" \
    ".Lsz11_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn11_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz21_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn21_%=\n\t" \
    ".Lsz21_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn21_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz31_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn31_%=\n\t" \
    ".Lsz31_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn31_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz41_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn41_%=\n\t" \
    ".Lsz41_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn41_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz51_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn51_%=\n\t" \
    ".Lsz51_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn51_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz61_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn61_%=\n\t" \
    ".Lsz61_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn61_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz71_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn71_%=\n\t" \
    ".Lsz71_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn71_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz81_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn81_%=\n\t" \
    ".Lsz81_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn81_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz91_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn91_%=\n\t" \
    ".Lsz91_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn91_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz101_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn101_%=\n\t" \
    ".Lsz101_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn101_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz111_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn111_%=\n\t" \
    ".Lsz111_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn111_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz121_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn121_%=\n\t" \
    ".Lsz121_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn121_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz131_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn131_%=\n\t" \
    ".Lsz131_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn131_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz141_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn141_%=\n\t" \
    ".Lsz141_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn141_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    ".Lgvp0s1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    ".Lgvp1s1_%=:\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    ".Lgvp2s1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    ".Lgvp3s1_%=:\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v112, v0\n\t" \
    "v_fma_f32 v21, s83, v113, v21\n\t" \
    "v_fma_f32 v42, s83, v114, v42\n\t" \
    "v_fma_f32 v63, s83, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    "v_mad_u32_u16 v117, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v117\n\t" \
    "ds_read_u16_d16_hi v112, v117\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v112, v0\n\t" \
    "v_fma_f32 v21, s99, v113, v21\n\t" \
    "v_fma_f32 v42, s99, v114, v42\n\t" \
    "v_fma_f32 v63, s99, v115, v63\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v116, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v117, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v116, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    ".Lgva1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v84, v0\n\t" \
    "v_fma_f32 v21, s69, v85, v21\n\t" \
    "v_fma_f32 v42, s69, v86, v42\n\t" \
    "v_fma_f32 v63, s69, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v88, v0\n\t" \
    "v_fma_f32 v21, s71, v89, v21\n\t" \
    "v_fma_f32 v42, s71, v90, v42\n\t" \
    "v_fma_f32 v63, s71, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v92, v0\n\t" \
    "v_fma_f32 v21, s73, v93, v21\n\t" \
    "v_fma_f32 v42, s73, v94, v42\n\t" \
    "v_fma_f32 v63, s73, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v96, v0\n\t" \
    "v_fma_f32 v21, s75, v97, v21\n\t" \
    "v_fma_f32 v42, s75, v98, v42\n\t" \
    "v_fma_f32 v63, s75, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v100, v0\n\t" \
    "v_fma_f32 v21, s77, v101, v21\n\t" \
    "v_fma_f32 v42, s77, v102, v42\n\t" \
    "v_fma_f32 v63, s77, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v104, v0\n\t" \
    "v_fma_f32 v21, s79, v105, v21\n\t" \
    "v_fma_f32 v42, s79, v106, v42\n\t" \
    "v_fma_f32 v63, s79, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v108, v0\n\t" \
    "v_fma_f32 v21, s81, v109, v21\n\t" \
    "v_fma_f32 v42, s81, v110, v42\n\t" \
    "v_fma_f32 v63, s81, v111, v63\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v116, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[86:87], v116\n\t" \
    "ds_read_u16_d16_hi v84, v116\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[90:91], v117\n\t" \
    "ds_read_u16_d16_hi v88, v117\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[94:95], v116\n\t" \
    "ds_read_u16_d16_hi v92, v116\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[98:99], v117\n\t" \
    "ds_read_u16_d16_hi v96, v117\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[102:103], v116\n\t" \
    "ds_read_u16_d16_hi v100, v116\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v117, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[106:107], v117\n\t" \
    "ds_read_u16_d16_hi v104, v117\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v116, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[110:111], v116\n\t" \
    "ds_read_u16_d16_hi v108, v116\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v85, 0xffff0000, v86\n\t" \
    "v_lshlrev_b32_e32 v86, 16, v87\n\t" \
    "v_and_b32_e32 v87, 0xffff0000, v87\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v89, 0xffff0000, v90\n\t" \
    "v_lshlrev_b32_e32 v90, 16, v91\n\t" \
    "v_and_b32_e32 v91, 0xffff0000, v91\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v93, 0xffff0000, v94\n\t" \
    "v_lshlrev_b32_e32 v94, 16, v95\n\t" \
    "v_and_b32_e32 v95, 0xffff0000, v95\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v97, 0xffff0000, v98\n\t" \
    "v_lshlrev_b32_e32 v98, 16, v99\n\t" \
    "v_and_b32_e32 v99, 0xffff0000, v99\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v101, 0xffff0000, v102\n\t" \
    "v_lshlrev_b32_e32 v102, 16, v103\n\t" \
    "v_and_b32_e32 v103, 0xffff0000, v103\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v105, 0xffff0000, v106\n\t" \
    "v_lshlrev_b32_e32 v106, 16, v107\n\t" \
    "v_and_b32_e32 v107, 0xffff0000, v107\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v109, 0xffff0000, v110\n\t" \
    "v_lshlrev_b32_e32 v110, 16, v111\n\t" \
    "v_and_b32_e32 v111, 0xffff0000, v111\n\t" \
    ".Lgvb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v84, v0\n\t" \
    "v_fma_f32 v21, s85, v85, v21\n\t" \
    "v_fma_f32 v42, s85, v86, v42\n\t" \
    "v_fma_f32 v63, s85, v87, v63\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v88, v0\n\t" \
    "v_fma_f32 v21, s87, v89, v21\n\t" \
    "v_fma_f32 v42, s87, v90, v42\n\t" \
    "v_fma_f32 v63, s87, v91, v63\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v92, v0\n\t" \
    "v_fma_f32 v21, s89, v93, v21\n\t" \
    "v_fma_f32 v42, s89, v94, v42\n\t" \
    "v_fma_f32 v63, s89, v95, v63\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v96, v0\n\t" \
    "v_fma_f32 v21, s91, v97, v21\n\t" \
    "v_fma_f32 v42, s91, v98, v42\n\t" \
    "v_fma_f32 v63, s91, v99, v63\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v100, v0\n\t" \
    "v_fma_f32 v21, s93, v101, v21\n\t" \
    "v_fma_f32 v42, s93, v102, v42\n\t" \
    "v_fma_f32 v63, s93, v103, v63\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v104, v0\n\t" \
    "v_fma_f32 v21, s95, v105, v21\n\t" \
    "v_fma_f32 v42, s95, v106, v42\n\t" \
    "v_fma_f32 v63, s95, v107, v63\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v108, v0\n\t" \
    "v_fma_f32 v21, s97, v109, v21\n\t" \
    "v_fma_f32 v42, s97, v110, v42\n\t" \
    "v_fma_f32 v63, s97, v111, v63\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_BF16_R21 \
    "v84", "v85", "v86", "v87", "v88", "v89", "v90", "v91", "v92", "v93", "v94", "v95", \
    "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103", "v104", "v105", "v106", "v107", \
    "v108", "v109", "v110", "v111", "v112", "v113", "v114", "v115", "v116", "v117", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"
// BF16 D, 28 rows per wave: accumulators v0..v111, component c of row slot q
// in v[28 c + q]; scratch v112..v145
#define GCOO_PANEL_SEGMENT_ASM_BF16_R28 \
    "s_mov_b32 s28, m0\n\t" \
    "v_mov_b32 v112, 0\n\t" \
    "v_mov_b32 v116, 0\n\t" \
    "v_mov_b32 v120, 0\n\t" \
    "v_mov_b32 v124, 0\n\t" \
    "v_mov_b32 v128, 0\n\t" \
    "v_mov_b32 v132, 0\n\t" \
    "v_mov_b32 v136, 0\n\t" \
    "v_mov_b32 v140, 0\n\t" \
    "s_nop 4\n\t" \
    "v_readlane_b32 s29, %[vpbo], %[sli]\n\t" \
    "v_readlane_b32 s30, %[vpn], %[sli]\n\t" \
    "s_cmp_eq_u32 %[par], 0\n\t" \
    "s_cbranch_scc0 .Lo1_%=\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lo1_%=:\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lc1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    "s_branch .Lc1_%=\n\t" \
    ".Lc0_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw0_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw0_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp0_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp0_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s24\n\t" \
    ".Lnp0_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd0_%=\n\t" \
    ".Lsg0_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd0_%=\n\t" \
    "s_add_u32 m0, %[sdst], 0x12c00\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz00_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn00_%=\n\t" \
    ".Lsz00_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn00_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz10_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn10_%=\n\t" \
    ".Lsz10_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn10_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz20_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn20_%=\n\t" \
    ".Lsz20_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn20_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz30_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn30_%=\n\t" \
    ".Lsz30_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn30_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz40_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn40_%=\n\t" \
    ".Lsz40_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn40_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz50_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn50_%=\n\t" \
    ".Lsz50_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn50_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz60_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn60_%=\n\t" \
    ".Lsz60_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn60_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz70_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn70_%=\n\t" \
    ".Lsz70_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn70_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz80_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn80_%=\n\t" \
    ".Lsz80_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn80_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz90_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn90_%=\n\t" \
    ".Lsz90_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn90_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz100_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn100_%=\n\t" \
    ".Lsz100_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn100_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz110_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn110_%=\n\t" \
    ".Lsz110_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn110_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz120_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn120_%=\n\t" \
    ".Lsz120_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn120_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz130_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn130_%=\n\t" \
    ".Lsz130_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn130_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd0_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz140_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn140_%=\n\t" \
    ".Lsz140_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn140_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd0_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpr20_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr0_%=\n\t" \
    ".Lpr20_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr0_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf00_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r0_%=\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    ".Lp0r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s0_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    ".Lgvp0s0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    ".Lggp0s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf00_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq00_%=\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf10_%=\n\t" \
    ".Lq00_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz00_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r0_%=\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    ".Lp1r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s0_%=\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    ".Lgvp1s0_%=:\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s0_%=\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    ".Lggp1s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz00_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf10_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq10_%=\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf20_%=\n\t" \
    ".Lq10_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz10_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r0_%=\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    ".Lp2r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s0_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    ".Lgvp2s0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    ".Lggp2s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz10_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf20_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq20_%=\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf30_%=\n\t" \
    ".Lq20_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz20_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r0_%=\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    ".Lp3r0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s0_%=\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    ".Lgvp3s0_%=:\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s0_%=\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    ".Lggp3s0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lz20_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lf30_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta0_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpa0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa0_%=\n\t" \
    ".Lpa0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa0_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s50, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna0_%=\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s26\n\t" \
    ".Lna0_%=:\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s44, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    "s_set_gpr_idx_idx s50\n\t" \
    "v_fma_f32 v0, s51, v140, v0\n\t" \
    "v_fma_f32 v28, s51, v141, v28\n\t" \
    "v_fma_f32 v56, s51, v142, v56\n\t" \
    "v_fma_f32 v84, s51, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb0_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb0_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpb0_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb0_%=\n\t" \
    ".Lpb0_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb0_%=:\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s66, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb0_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s26\n\t" \
    ".Lnb0_%=:\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s60, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    "s_set_gpr_idx_idx s66\n\t" \
    "v_fma_f32 v0, s67, v140, v0\n\t" \
    "v_fma_f32 v28, s67, v141, v28\n\t" \
    "v_fma_f32 v56, s67, v142, v56\n\t" \
    "v_fma_f32 v84, s67, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce0_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba0_%=\n\t" \
    ".Lta0_%=:\n\t" \
    "v_mad_u32_u16 v144, s36, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s38, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s40, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s42, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s44, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v145, s46, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa0_%=\n\t" \
    "v_mad_u32_u16 v144, s48, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    ".Lgwa0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva0_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    ".Lgva0_%=:\n\t" \
    "s_set_gpr_idx_on s36, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s37, v112, v0\n\t" \
    "v_fma_f32 v28, s37, v113, v28\n\t" \
    "v_fma_f32 v56, s37, v114, v56\n\t" \
    "v_fma_f32 v84, s37, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s38\n\t" \
    "v_fma_f32 v0, s39, v116, v0\n\t" \
    "v_fma_f32 v28, s39, v117, v28\n\t" \
    "v_fma_f32 v56, s39, v118, v56\n\t" \
    "v_fma_f32 v84, s39, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s40\n\t" \
    "v_fma_f32 v0, s41, v120, v0\n\t" \
    "v_fma_f32 v28, s41, v121, v28\n\t" \
    "v_fma_f32 v56, s41, v122, v56\n\t" \
    "v_fma_f32 v84, s41, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s42\n\t" \
    "v_fma_f32 v0, s43, v124, v0\n\t" \
    "v_fma_f32 v28, s43, v125, v28\n\t" \
    "v_fma_f32 v56, s43, v126, v56\n\t" \
    "v_fma_f32 v84, s43, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s44\n\t" \
    "v_fma_f32 v0, s45, v128, v0\n\t" \
    "v_fma_f32 v28, s45, v129, v28\n\t" \
    "v_fma_f32 v56, s45, v130, v56\n\t" \
    "v_fma_f32 v84, s45, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s46\n\t" \
    "v_fma_f32 v0, s47, v132, v0\n\t" \
    "v_fma_f32 v28, s47, v133, v28\n\t" \
    "v_fma_f32 v56, s47, v134, v56\n\t" \
    "v_fma_f32 v84, s47, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea0_%=\n\t" \
    "s_set_gpr_idx_idx s48\n\t" \
    "v_fma_f32 v0, s49, v136, v0\n\t" \
    "v_fma_f32 v28, s49, v137, v28\n\t" \
    "v_fma_f32 v56, s49, v138, v56\n\t" \
    "v_fma_f32 v84, s49, v139, v84\n\t" \
    ".Lgea0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Ltb0_%=:\n\t" \
    "v_mad_u32_u16 v144, s52, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s54, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s56, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s58, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s60, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v145, s62, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb0_%=\n\t" \
    "v_mad_u32_u16 v144, s64, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    ".Lgwb0_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb0_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    ".Lgvb0_%=:\n\t" \
    "s_set_gpr_idx_on s52, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s53, v112, v0\n\t" \
    "v_fma_f32 v28, s53, v113, v28\n\t" \
    "v_fma_f32 v56, s53, v114, v56\n\t" \
    "v_fma_f32 v84, s53, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s54\n\t" \
    "v_fma_f32 v0, s55, v116, v0\n\t" \
    "v_fma_f32 v28, s55, v117, v28\n\t" \
    "v_fma_f32 v56, s55, v118, v56\n\t" \
    "v_fma_f32 v84, s55, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s56\n\t" \
    "v_fma_f32 v0, s57, v120, v0\n\t" \
    "v_fma_f32 v28, s57, v121, v28\n\t" \
    "v_fma_f32 v56, s57, v122, v56\n\t" \
    "v_fma_f32 v84, s57, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s58\n\t" \
    "v_fma_f32 v0, s59, v124, v0\n\t" \
    "v_fma_f32 v28, s59, v125, v28\n\t" \
    "v_fma_f32 v56, s59, v126, v56\n\t" \
    "v_fma_f32 v84, s59, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s60\n\t" \
    "v_fma_f32 v0, s61, v128, v0\n\t" \
    "v_fma_f32 v28, s61, v129, v28\n\t" \
    "v_fma_f32 v56, s61, v130, v56\n\t" \
    "v_fma_f32 v84, s61, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s62\n\t" \
    "v_fma_f32 v0, s63, v132, v0\n\t" \
    "v_fma_f32 v28, s63, v133, v28\n\t" \
    "v_fma_f32 v56, s63, v134, v56\n\t" \
    "v_fma_f32 v84, s63, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb0_%=\n\t" \
    "s_set_gpr_idx_idx s64\n\t" \
    "v_fma_f32 v0, s65, v136, v0\n\t" \
    "v_fma_f32 v28, s65, v137, v28\n\t" \
    "v_fma_f32 v56, s65, v138, v56\n\t" \
    "v_fma_f32 v84, s65, v139, v84\n\t" \
    ".Lgeb0_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce0_%=\n\t" \
    ".Lce0_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    ".Lc1_%=:\n\t" \
    "s_mov_b32 s25, s30\n\t" \
    "s_add_u32 s26, s29, 0x80\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lw1_%=\n\t" \
    "s_waitcnt vmcnt(0)\n\t" \
    ".Lw1_%=:\n\t" \
    "s_barrier\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_add_u32 s24, %[sli], 1\n\t" \
    "s_cmp_lt_u32 s24, %[liend]\n\t" \
    "s_cbranch_scc0 .Lnp1_%=\n\t" \
    "v_readlane_b32 s29, %[vpbo], s24\n\t" \
    "v_readlane_b32 s30, %[vpn], s24\n\t" \
    "s_cmp_eq_u32 s30, 0\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_load_dwordx16 s[36:51], %[pbase], s29\n\t" \
    "s_cmp_le_u32 s30, 8\n\t" \
    "s_cbranch_scc1 .Lnp1_%=\n\t" \
    "s_add_u32 s24, s29, 0x40\n\t" \
    "s_load_dwordx16 s[52:67], %[pbase], s24\n\t" \
    ".Lnp1_%=:\n\t" \
    "s_cmp_le_u32 %[sso], %[lim]\n\t" \
    "s_cbranch_scc0 .Lsg1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsd1_%=\n\t" \
    ".Lsg1_%=:\n\t" \
    "s_cmp_eq_u32 %[kb], 0\n\t" \
    "s_cbranch_scc1 .Lsd1_%=\n\t" \
    "s_mov_b32 m0, %[sdst]\n\t" \
    "s_mov_b32 s24, %[sso]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz01_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn01_%=\n\t" \
    ".Lsz01_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn01_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz11_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn11_%=\n\t" \
    ".Lsz11_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn11_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz21_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn21_%=\n\t" \
    ".Lsz21_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn21_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz31_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn31_%=\n\t" \
    ".Lsz31_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn31_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz41_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn41_%=\n\t" \
    ".Lsz41_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn41_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz51_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn51_%=\n\t" \
    ".Lsz51_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn51_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz61_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn61_%=\n\t" \
    ".Lsz61_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn61_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz71_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn71_%=\n\t" \
    ".Lsz71_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn71_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz81_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn81_%=\n\t" \
    ".Lsz81_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn81_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz91_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn91_%=\n\t" \
    ".Lsz91_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn91_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz101_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn101_%=\n\t" \
    ".Lsz101_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn101_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz111_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn111_%=\n\t" \
    ".Lsz111_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn111_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz121_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn121_%=\n\t" \
    ".Lsz121_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn121_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz131_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn131_%=\n\t" \
    ".Lsz131_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn131_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    "s_cmp_lt_u32 s24, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsd1_%=\n\t" \
    "s_add_u32 s27, s24, %[row1]\n\t" \
    "s_cmp_lt_u32 s27, %[kb]\n\t" \
    "s_cbranch_scc0 .Lsz141_%=\n\t" \
    "buffer_load_dwordx4 %[voff], %[rsrc], s24 offen lds\n\t" \
    "s_branch .Lsn141_%=\n\t" \
    ".Lsz141_%=:\n\t" \
    "buffer_load_dwordx4 %[voffz], %[rsrc], s24 offen lds\n\t" \
    ".Lsn141_%=:\n\t" \
    "s_add_u32 m0, m0, 0x400\n\t" \
    "s_add_u32 s24, s24, %[rowb]\n\t" \
    ".Lsd1_%=:\n\t" \
    "s_add_u32 %[sso], %[sso], %[tkrowb]\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_cmp_ge_u32 s25, 21\n\t" \
    "s_cbranch_scc0 .Lpr21_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lpr1_%=\n\t" \
    ".Lpr21_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lpr1_%=:\n\t" \
    "s_cmp_ge_u32 s25, 4\n\t" \
    "s_cbranch_scc1 .Lf01_%=\n\t" \
    "s_mov_b32 s27, s25\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp0r1_%=\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    ".Lp0r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp0s1_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    ".Lgvp0s1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp0s1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    ".Lggp0s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf01_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_ge_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lq01_%=\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf11_%=\n\t" \
    ".Lq01_%=:\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lz01_%=\n\t" \
    "s_sub_u32 s27, s25, 4\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp1r1_%=\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    ".Lp1r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp1s1_%=\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    ".Lgvp1s1_%=:\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp1s1_%=\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    ".Lggp1s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz01_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf11_%=:\n\t" \
    "s_cmp_ge_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lq11_%=\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf21_%=\n\t" \
    ".Lq11_%=:\n\t" \
    "s_cmp_gt_u32 s25, 8\n\t" \
    "s_cbranch_scc0 .Lz11_%=\n\t" \
    "s_sub_u32 s27, s25, 8\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp2r1_%=\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    ".Lp2r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp2s1_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    ".Lgvp2s1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp2s1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    ".Lggp2s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz11_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf21_%=:\n\t" \
    "s_cmp_ge_u32 s25, 16\n\t" \
    "s_cbranch_scc0 .Lq21_%=\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(8)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lf31_%=\n\t" \
    ".Lq21_%=:\n\t" \
    "s_cmp_gt_u32 s25, 12\n\t" \
    "s_cbranch_scc0 .Lz21_%=\n\t" \
    "s_sub_u32 s27, s25, 12\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lp3r1_%=\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    ".Lp3r1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lgvp3s1_%=\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    ".Lgvp3s1_%=:\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_cmp_gt_u32 s27, 1\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_cmp_gt_u32 s27, 2\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_cmp_gt_u32 s27, 3\n\t" \
    "s_cbranch_scc0 .Lggp3s1_%=\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    ".Lggp3s1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lz21_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lf31_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_cmp_le_u32 s25, 16\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_sub_u32 s25, s25, 16\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    "s_add_u32 s26, s26, 0x40\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lba1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lta1_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpa1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqa1_%=\n\t" \
    ".Lpa1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqa1_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s82, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lna1_%=\n\t" \
    "s_load_dwordx16 s[84:99], %[pbase], s26\n\t" \
    ".Lna1_%=:\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s76, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    "s_set_gpr_idx_idx s82\n\t" \
    "v_fma_f32 v0, s83, v140, v0\n\t" \
    "v_fma_f32 v28, s83, v141, v28\n\t" \
    "v_fma_f32 v56, s83, v142, v56\n\t" \
    "v_fma_f32 v84, s83, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    ".Lbb1_%=:\n\t" \
    "s_cmp_lt_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Ltb1_%=\n\t" \
    "s_cmp_ge_u32 s25, 28\n\t" \
    "s_cbranch_scc0 .Lpb1_%=\n\t" \
    "s_setprio 3\n\t" \
    "s_branch .Lqb1_%=\n\t" \
    ".Lpb1_%=:\n\t" \
    "s_setprio 2\n\t" \
    ".Lqb1_%=:\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    "v_mad_u32_u16 v145, s98, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[142:143], v145\n\t" \
    "ds_read_u16_d16_hi v140, v145\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_cmp_le_u32 s25, 8\n\t" \
    "s_cbranch_scc1 .Lnb1_%=\n\t" \
    "s_load_dwordx16 s[68:83], %[pbase], s26\n\t" \
    ".Lnb1_%=:\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    "v_and_b32_e32 v141, 0xffff0000, v142\n\t" \
    "v_lshlrev_b32_e32 v142, 16, v143\n\t" \
    "v_and_b32_e32 v143, 0xffff0000, v143\n\t" \
    "s_set_gpr_idx_on s92, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    "s_set_gpr_idx_idx s98\n\t" \
    "v_fma_f32 v0, s99, v140, v0\n\t" \
    "v_fma_f32 v28, s99, v141, v28\n\t" \
    "v_fma_f32 v56, s99, v142, v56\n\t" \
    "v_fma_f32 v84, s99, v143, v84\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_sub_u32 s25, s25, 8\n\t" \
    "s_add_u32 s26, s26, 64\n\t" \
    "s_cmp_eq_u32 s25, 0\n\t" \
    "s_cbranch_scc1 .Lce1_%=\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "s_branch .Lba1_%=\n\t" \
    ".Lta1_%=:\n\t" \
    "v_mad_u32_u16 v144, s68, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s70, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s72, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s74, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s76, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v145, s78, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwa1_%=\n\t" \
    "v_mad_u32_u16 v144, s80, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    ".Lgwa1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgva1_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    ".Lgva1_%=:\n\t" \
    "s_set_gpr_idx_on s68, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s69, v112, v0\n\t" \
    "v_fma_f32 v28, s69, v113, v28\n\t" \
    "v_fma_f32 v56, s69, v114, v56\n\t" \
    "v_fma_f32 v84, s69, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s70\n\t" \
    "v_fma_f32 v0, s71, v116, v0\n\t" \
    "v_fma_f32 v28, s71, v117, v28\n\t" \
    "v_fma_f32 v56, s71, v118, v56\n\t" \
    "v_fma_f32 v84, s71, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s72\n\t" \
    "v_fma_f32 v0, s73, v120, v0\n\t" \
    "v_fma_f32 v28, s73, v121, v28\n\t" \
    "v_fma_f32 v56, s73, v122, v56\n\t" \
    "v_fma_f32 v84, s73, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s74\n\t" \
    "v_fma_f32 v0, s75, v124, v0\n\t" \
    "v_fma_f32 v28, s75, v125, v28\n\t" \
    "v_fma_f32 v56, s75, v126, v56\n\t" \
    "v_fma_f32 v84, s75, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s76\n\t" \
    "v_fma_f32 v0, s77, v128, v0\n\t" \
    "v_fma_f32 v28, s77, v129, v28\n\t" \
    "v_fma_f32 v56, s77, v130, v56\n\t" \
    "v_fma_f32 v84, s77, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s78\n\t" \
    "v_fma_f32 v0, s79, v132, v0\n\t" \
    "v_fma_f32 v28, s79, v133, v28\n\t" \
    "v_fma_f32 v56, s79, v134, v56\n\t" \
    "v_fma_f32 v84, s79, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgea1_%=\n\t" \
    "s_set_gpr_idx_idx s80\n\t" \
    "v_fma_f32 v0, s81, v136, v0\n\t" \
    "v_fma_f32 v28, s81, v137, v28\n\t" \
    "v_fma_f32 v56, s81, v138, v56\n\t" \
    "v_fma_f32 v84, s81, v139, v84\n\t" \
    ".Lgea1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Ltb1_%=:\n\t" \
    "v_mad_u32_u16 v144, s84, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[114:115], v144\n\t" \
    "ds_read_u16_d16_hi v112, v144\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s86, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[118:119], v145\n\t" \
    "ds_read_u16_d16_hi v116, v145\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s88, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[122:123], v144\n\t" \
    "ds_read_u16_d16_hi v120, v144\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s90, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[126:127], v145\n\t" \
    "ds_read_u16_d16_hi v124, v145\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s92, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[130:131], v144\n\t" \
    "ds_read_u16_d16_hi v128, v144\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v145, s94, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[134:135], v145\n\t" \
    "ds_read_u16_d16_hi v132, v145\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgwb1_%=\n\t" \
    "v_mad_u32_u16 v144, s96, 16, %[vb0] op_sel:[1,0,0,0]\n\t" \
    "ds_read_b64 v[138:139], v144\n\t" \
    "ds_read_u16_d16_hi v136, v144\n\t" \
    ".Lgwb1_%=:\n\t" \
    "s_waitcnt lgkmcnt(0)\n\t" \
    "v_and_b32_e32 v113, 0xffff0000, v114\n\t" \
    "v_lshlrev_b32_e32 v114, 16, v115\n\t" \
    "v_and_b32_e32 v115, 0xffff0000, v115\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v117, 0xffff0000, v118\n\t" \
    "v_lshlrev_b32_e32 v118, 16, v119\n\t" \
    "v_and_b32_e32 v119, 0xffff0000, v119\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v121, 0xffff0000, v122\n\t" \
    "v_lshlrev_b32_e32 v122, 16, v123\n\t" \
    "v_and_b32_e32 v123, 0xffff0000, v123\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v125, 0xffff0000, v126\n\t" \
    "v_lshlrev_b32_e32 v126, 16, v127\n\t" \
    "v_and_b32_e32 v127, 0xffff0000, v127\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v129, 0xffff0000, v130\n\t" \
    "v_lshlrev_b32_e32 v130, 16, v131\n\t" \
    "v_and_b32_e32 v131, 0xffff0000, v131\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v133, 0xffff0000, v134\n\t" \
    "v_lshlrev_b32_e32 v134, 16, v135\n\t" \
    "v_and_b32_e32 v135, 0xffff0000, v135\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgvb1_%=\n\t" \
    "v_and_b32_e32 v137, 0xffff0000, v138\n\t" \
    "v_lshlrev_b32_e32 v138, 16, v139\n\t" \
    "v_and_b32_e32 v139, 0xffff0000, v139\n\t" \
    ".Lgvb1_%=:\n\t" \
    "s_set_gpr_idx_on s84, gpr_idx(SRC2,DST)\n\t" \
    "v_fma_f32 v0, s85, v112, v0\n\t" \
    "v_fma_f32 v28, s85, v113, v28\n\t" \
    "v_fma_f32 v56, s85, v114, v56\n\t" \
    "v_fma_f32 v84, s85, v115, v84\n\t" \
    "s_cmp_gt_u32 s25, 1\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s86\n\t" \
    "v_fma_f32 v0, s87, v116, v0\n\t" \
    "v_fma_f32 v28, s87, v117, v28\n\t" \
    "v_fma_f32 v56, s87, v118, v56\n\t" \
    "v_fma_f32 v84, s87, v119, v84\n\t" \
    "s_cmp_gt_u32 s25, 2\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s88\n\t" \
    "v_fma_f32 v0, s89, v120, v0\n\t" \
    "v_fma_f32 v28, s89, v121, v28\n\t" \
    "v_fma_f32 v56, s89, v122, v56\n\t" \
    "v_fma_f32 v84, s89, v123, v84\n\t" \
    "s_cmp_gt_u32 s25, 3\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s90\n\t" \
    "v_fma_f32 v0, s91, v124, v0\n\t" \
    "v_fma_f32 v28, s91, v125, v28\n\t" \
    "v_fma_f32 v56, s91, v126, v56\n\t" \
    "v_fma_f32 v84, s91, v127, v84\n\t" \
    "s_cmp_gt_u32 s25, 4\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s92\n\t" \
    "v_fma_f32 v0, s93, v128, v0\n\t" \
    "v_fma_f32 v28, s93, v129, v28\n\t" \
    "v_fma_f32 v56, s93, v130, v56\n\t" \
    "v_fma_f32 v84, s93, v131, v84\n\t" \
    "s_cmp_gt_u32 s25, 5\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s94\n\t" \
    "v_fma_f32 v0, s95, v132, v0\n\t" \
    "v_fma_f32 v28, s95, v133, v28\n\t" \
    "v_fma_f32 v56, s95, v134, v56\n\t" \
    "v_fma_f32 v84, s95, v135, v84\n\t" \
    "s_cmp_gt_u32 s25, 6\n\t" \
    "s_cbranch_scc0 .Lgeb1_%=\n\t" \
    "s_set_gpr_idx_idx s96\n\t" \
    "v_fma_f32 v0, s97, v136, v0\n\t" \
    "v_fma_f32 v28, s97, v137, v28\n\t" \
    "v_fma_f32 v56, s97, v138, v56\n\t" \
    "v_fma_f32 v84, s97, v139, v84\n\t" \
    ".Lgeb1_%=:\n\t" \
    "s_set_gpr_idx_off\n\t" \
    "s_branch .Lce1_%=\n\t" \
    ".Lce1_%=:\n\t" \
    "s_setprio 0\n\t" \
    "s_add_u32 %[sli], %[sli], 1\n\t" \
    "s_cmp_lt_u32 %[sli], %[liend]\n\t" \
    "s_cbranch_scc0 .Lx_%=\n\t" \
    "s_branch .Lc0_%=\n\t" \
    ".Lx_%=:\n\t" \
    "s_mov_b32 m0, s28\n\t"
#define GCOO_PANEL_SEGMENT_CLOBBERS_BF16_R28 \
    "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", "v122", "v123", \
    "v124", "v125", "v126", "v127", "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135", \
    "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143", "v144", "v145", "s24", "s25", \
    "s26", "s27", "s28", "s29", "s30", "s36", "s37", "s38", "s39", "s40", "s41", "s42", \
    "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54", \
    "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", \
    "s67", "s68", "s69", "s70", "s71", "s72", "s73", "s74", "s75", "s76", "s77", "s78", \
    "s79", "s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90", \
    "s91", "s92", "s93", "s94", "s95", "s96", "s97", "s98", "s99", "scc", "memory"

// Host-side launchers of the HIP kernels (one declaration per phase).
#pragma once
#include "internal.h"

namespace dclust {

// start of a run: control counters, statistics, radix pass flags, bbox
void launch_init_head(const Dev& d, hipStream_t s);
void launch_readback(const Dev& d, long long* out, hipStream_t s);
// per-run state of the clustering phases (labels, forest, minima, flags,
// per-cell core counts, per-wave pair counts)
void launch_init_run(const Dev& d, int64_t waves, hipStream_t s);
// bbox + grid set-up (per-instance cell width, cells per axis, key space)
void launch_index_head(const Dev& d, int nchunks, hipStream_t s);
// the index after the grid set-up, in three parts: cell keys, `passes`
// radix passes and CSR compaction (tile size from the CU count `cus`); the
// gather of the coordinates; the line table (the last two are independent)
void launch_index_sort(const Dev& d, int64_t n, int nchunks, int passes, int cus, hipStream_t s);
void launch_gather(const Dev& d, int64_t n, hipStream_t s);
void launch_lines(const Dev& d, int grid_cells, hipStream_t s);
// neighbour rows of every cell (k_rows)
void launch_rows(const Dev& d, int64_t n, int grid_cells, hipStream_t s);
// exclusive in-place scan of nb int64 values by one block
void launch_scan_top(int64_t* v, int nb, hipStream_t s);

int64_t sort_hist_entries(int64_t n);
// per-stripe capacity of the row pass's work list for up to n cells
int64_t row_list_cap(int64_t n);
int64_t scan_tmp_entries(int64_t n);

void launch_core(const Dev& d, bool f32, hipStream_t s);
void launch_adj(const Dev& d, bool f32, int heavy_grid, hipStream_t s);
// its two kernels apart: light cells (k_adj) and heavy cells (k_adj_heavy)
// are independent (both only OR bits into the masks)
void launch_adj_light(const Dev& d, bool f32, hipStream_t s);
void launch_adj_heavy(const Dev& d, bool f32, int heavy_grid, hipStream_t s);
void launch_pass_prep(const Dev& d, int pass, hipStream_t s);
void launch_pass_carry(const Dev& d, int pass, int check, hipStream_t s);
// after `passes` expansion passes: pending work abandons the rest of the run
void launch_expand(const Dev& d, int pass, int grid, int seeds, int threads, int budget,
                   hipStream_t s);
void launch_merge(const Dev& d, int grid, hipStream_t s);
// workgroups of the producers of the striped work lists (their capacity)
int64_t point_blocks(int64_t n);
int64_t cell_comp_blocks(int64_t n);
// far stencil pairs across components, then final roots
void launch_far(const Dev& d, bool f32, int grid, hipStream_t s);
void launch_border(const Dev& d, bool f32, hipStream_t s);
void launch_relabel(const Dev& d, int64_t n, hipStream_t s);
// core flags of the last run (cell-sorted core[], orig[]) to input order
void launch_core_out(const uint8_t* core, const int32_t* orig, int64_t n, uint8_t* out,
                     hipStream_t s);

// one-step baseline: state reset, refill seeds (first[b] < 0: none), one
// expansion step of every chain (chain ids chain0 + b), final border pass
void launch_os_init(const OneStep& o, int64_t n, hipStream_t s);
void launch_os_refill(const OneStep& o, const int32_t* first, hipStream_t s);
void launch_os_step(const Dev& d, const OneStep& o, int32_t chain0, int threads,
                    hipStream_t s);
void launch_os_border(const Dev& d, const OneStep& o, hipStream_t s);

// occupancy helper for the expansion kernel
int expand_blocks_per_cu(int seeds, int threads);

}  // namespace dclust

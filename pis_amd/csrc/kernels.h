// kernels.h -- host-side launch wrappers of the gfx950 kernels (kernels.hip).
#pragma once
#include "common.h"

namespace pis {

void launch_load_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, const double* pos,
                     const double* vel, const double* force, const uint8_t* type);
void launch_store_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, double* pos,
                      double* vel, double* force);
void launch_set_ctl(hipStream_t st, const SysDev* d_sys, int nsys, int slot, int value);

void launch_integrate_pos(hipStream_t st, const SysDev* d_sys, const WorkItem* atom_work,
                          int n_atom_work);
// Counting sort of every system whose rebuild flag is set (bin, scan, scatter, rank).
void launch_rebuild_sort(hipStream_t st, const SysDev* d_sys, int nsys, const WorkItem* atom_work,
                         int n_atom_work);
// Neighbour lists of the tiles of one staging-capacity class.
void launch_build_list(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work, int n_tile_work,
                       int scap, int kinds);
void launch_force(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work,
                  int n_tile_work, int scap, bool mixed, int integrate, int kinds);
// `step` = 1 after an NVE step: frozen systems (ctl[CTL_LEFT] == 0) are skipped and a
// positive step budget is counted down.
void launch_reduce(hipStream_t st, const SysDev* d_sys, int nsys, int flip, int record, int step);
void launch_gather_ctl(hipStream_t st, const SysDev* d_sys, int nsys, int32_t* out);
// NVT: start-of-step chain update and velocity scale; reset of the run's step counter.
void launch_nhc_pre(hipStream_t st, const SysDev* d_sys, int nsys);
void launch_nhc_run(hipStream_t st, const SysDev* d_sys, int nsys, uint64_t step0, uint64_t total);
void launch_flip(hipStream_t st, const SysDev* d_sys, int nsys);
// copy the live set's types and original indices into the other set (after a host sort)
void launch_sync_labels(hipStream_t st, const SysDev* d_sys, const WorkItem* atom_work, int n_atom_work);
// NPT around one NVT step: what = 0 start-of-step tensor sums, 1 barostat + box update and
// scaled state, 2 end-of-step scaling, momentum update and thermo record (record != 0).
void launch_npt(hipStream_t st, const SysDev* d_sys, int nsys, const WorkItem* atom_work, int n_atom_work,
                int what, int record);
void launch_list_to_orig(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work,
                         int n_tile_work, int scap, uint32_t* out, uint32_t* counts);

// Per atom (original index): list length and sum(mix64(original index)) of its entries.
void launch_list_digest(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work, int n_tile_work,
                        int scap, uint32_t* counts, uint64_t* hashes);

void launch_copy_calibration(hipStream_t st, const void* src, void* dst, size_t bytes);

}  // namespace pis

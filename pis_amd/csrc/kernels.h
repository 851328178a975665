// kernels.h -- host-side launch wrappers of the gfx950 kernels (kernels.hip).
#pragma once
#include "common.h"

namespace pis {

// Dynamic LDS bytes the tile kernels need for a staging capacity of `scap` atoms.
size_t tile_lds_bytes(int scap, bool multi_type);

void launch_load_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, const double* pos,
                     const double* vel, const double* force, const uint8_t* type);
void launch_store_aos(hipStream_t st, const SysDev* d_sys, int sys, int n, double* pos,
                      double* vel, double* force);
void launch_set_ctl(hipStream_t st, const SysDev* d_sys, int nsys, int slot, int value);

void launch_integrate_pos(hipStream_t st, const SysDev* d_sys, const WorkItem* atom_work,
                          int n_atom_work);
void launch_rebuild(hipStream_t st, const SysDev* d_sys, int nsys, const WorkItem* atom_work,
                    int n_atom_work, const WorkItem* tile_work, int n_tile_work, int scap,
                    bool multi_type);
void launch_force(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work,
                  int n_tile_work, int scap, bool multi_type, bool mixed, int integrate);
void launch_reduce(hipStream_t st, const SysDev* d_sys, int nsys, int flip, int record);
void launch_flip(hipStream_t st, const SysDev* d_sys, int nsys);
void launch_list_to_orig(hipStream_t st, const SysDev* d_sys, const WorkItem* tile_work,
                         int n_tile_work, int scap, bool multi_type, uint32_t* out,
                         uint32_t* counts);

void launch_copy_calibration(hipStream_t st, const void* src, void* dst, size_t bytes);

}  // namespace pis

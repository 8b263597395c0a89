// gmres.h — the batched GMRES kernel's shared pattern tables and launchers
// (gmres.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spx {

constexpr int kGmresThreads = 256;

struct DevGmres {
  int n, num_pos;
  const int* row_ptr;   // n + 1 over pattern positions, sorted by (row, col)
  const int* col;       // num_pos
  const int* pos_off;   // num_pos + 1 into pos_trip
  const int* pos_trip;  // triplet indices summed into each position, in order
  const int* diag_pos;  // n: position of (i, i), -1 if absent
};

// precision: 0 FP64 throughout, 1 FP32 Krylov solver inside FP64 refinement
size_t gmres_smem_bytes(int n, int restart, int precision);
int64_t gmres_slot_doubles(int n, int num_pos, int restart, int precision);
hipError_t gmres_grid(int n, int restart, int precision, int64_t B, int* grid);
hipError_t launch_gmres(const DevGmres& p, int64_t nnz, const double* values, const double* rhs,
                        const double* x0, int64_t B, int restart, int max_iter, double tol,
                        int jacobi, int precision, double* x, int* iters, double* res,
                        uint8_t* conv, double* ws, int grid, hipStream_t st);

}  // namespace spx

// FP64 CSR matrix utilities on the device: SpMV and row zeroing.
//
// Replace the reference's PETSc MatMult and Mat.zero_rows
// (pyop2/petsc_base.py:131-135).  The CG solver is in cg_batch.hip.
#include "common.h"

namespace {

constexpr int kBlock = 256;

// y = A x with G lanes per row (G a power of two <= 64), grid-stride over rows.
template <int G>
__global__ void __launch_bounds__(kBlock) spmv_kernel(int n, const int* __restrict__ rowptr,
                                                      const int* __restrict__ colidx,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y) {
  const int sub = threadIdx.x & (G - 1);
  const long long groups = static_cast<long long>(gridDim.x) * (kBlock / G);
  for (long long row = (static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x) / G; row < n;
       row += groups) {
    const int b = rowptr[row], e = rowptr[row + 1];
    double s = 0.0;
    for (int k = b + sub; k < e; k += G) s += vals[k] * x[colidx[k]];
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (sub == 0) y[row] = s;
  }
}

template <int G>
void launch(int grid, hipStream_t s, int n, const int* rp, const int* ci, const double* v,
            const double* x, double* y) {
  hipLaunchKernelGGL(spmv_kernel<G>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, y);
}

__global__ void zero_rows_kernel(int nrows_list, const int* __restrict__ rows,
                                 const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                 double* __restrict__ vals, double diag) {
  const int i = blockIdx.x;
  if (i >= nrows_list) return;
  const int r = rows[i];
  for (int k = rowptr[r] + threadIdx.x; k < rowptr[r + 1]; k += blockDim.x)
    vals[k] = colidx[k] == r ? diag : 0.0;
}

}  // namespace

extern "C" {

// y = A x for an FP64 CSR matrix with n rows.
int op2hip_spmv(int n, long long nnz, const int* rowptr, const int* colidx, const double* vals,
                const double* x, double* y, void* stream) {
  OP2HIP_REQUIRE(n >= 0, "bad spmv size");
  if (n == 0) return OP2HIP_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double avg = static_cast<double>(nnz) / n;
  int g = 1;
  while (g < 64 && g * 2 <= avg) g *= 2;
  long long grid = (static_cast<long long>(n) * g + kBlock - 1) / kBlock;
  if (grid > 65536) grid = 65536;
  const int gi = static_cast<int>(grid);
  switch (g) {
    case 1: launch<1>(gi, s, n, rowptr, colidx, vals, x, y); break;
    case 2: launch<2>(gi, s, n, rowptr, colidx, vals, x, y); break;
    case 4: launch<4>(gi, s, n, rowptr, colidx, vals, x, y); break;
    case 8: launch<8>(gi, s, n, rowptr, colidx, vals, x, y); break;
    case 16: launch<16>(gi, s, n, rowptr, colidx, vals, x, y); break;
    case 32: launch<32>(gi, s, n, rowptr, colidx, vals, x, y); break;
    default: launch<64>(gi, s, n, rowptr, colidx, vals, x, y); break;
  }
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// Zero the listed rows (device int32 row ids) except their diagonal, which
// is set to `diag` (PETSc MatZeroRowsLocal with KEEP_NONZERO_PATTERN).
int op2hip_csr_zero_rows(int nlist, const int* rows, const int* rowptr, const int* colidx,
                         double* vals, double diag, void* stream) {
  if (nlist <= 0) return OP2HIP_OK;
  hipLaunchKernelGGL(zero_rows_kernel, dim3(nlist), dim3(64), 0,
                     static_cast<hipStream_t>(stream), nlist, rows, rowptr, colidx, vals, diag);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

}  // extern "C"

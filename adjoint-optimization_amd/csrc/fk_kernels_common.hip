// fk_kernels_common.hip -- small kernels shared by both paths: objective and terminal adjoint,
// fixed-order row sums of FP64 partials, packing between the packed [u;v;w] layout of the API
// and the padded device planes, and the stimulus helper.
#include "fk_internal.hpp"

namespace fk {

// One thread per packed row r: (S y_v)_r in FP64 (row entries in increasing column order, the
// order of the reference's CSC scatter, G:1514-1522), z_v = -gamma_v (S y_v) written to the
// cell's plane entry, and the block partial of L0 = -sum_v y_v . z_v (G:1769-1802).
template <typename T>
__global__ __launch_bounds__(256) void k_objective(int n_num, const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ colind,
                                                   const double *__restrict__ val, double s_diag,
                                                   const int32_t *__restrict__ cell_of, int n,
                                                   const T *__restrict__ y, const double *__restrict__ gamma3,
                                                   T *__restrict__ z, double *__restrict__ L0_part)
{
    int r = blockIdx.x * 256 + threadIdx.x;
    double contrib = 0.0;
    if (r < n_num) {
        int cr = cell_of[r];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const T *yv = y + (size_t)v * n;
            double acc;
            if (rowptr) {
                acc = 0.0;
                for (int64_t q = rowptr[r]; q < rowptr[r + 1]; q++) acc += val[q] * (double)yv[cell_of[colind[q]]];
            } else {
                acc = s_diag * (double)yv[cr];
            }
            double zr = acc * -gamma3[v];
            z[(size_t)v * n + cr] = (T)zr;
            contrib -= (double)yv[cr] * zr;
        }
    }
    double s = block_sum_256(contrib, threadIdx.x);
    if (threadIdx.x == 0) L0_part[blockIdx.x] = s;
}

int objective_blocks(int n_num) { return (n_num + 255) / 256; }

template <typename T>
void objective_launch(int n_num, const int64_t *rowptr, const int32_t *colind, const double *val, double s_diag,
                      const int32_t *cell_of, int n, const T *y, const double *gamma3, T *z, double *L0_part,
                      hipStream_t stream)
{
    hipLaunchKernelGGL(k_objective<T>, dim3(objective_blocks(n_num)), dim3(256), 0, stream, n_num, rowptr, colind,
                       val, s_diag, cell_of, n, y, gamma3, z, L0_part);
}

__global__ void k_sum_rows(const double *__restrict__ part, int nb, int nrows_out, double *__restrict__ out)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows_out) return;
    const double *p = part + (size_t)r * nb;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += p[b];
    out[r] = s;
}

void sum_rows_launch(const double *part, int nb, int nrows_out, double *out, hipStream_t stream)
{
    if (nrows_out <= 0) return;
    hipLaunchKernelGGL(k_sum_rows, dim3((nrows_out + 255) / 256), dim3(256), 0, stream, part, nb, nrows_out, out);
}

template <typename T>
__global__ void k_pack(const T *__restrict__ planes, const int32_t *__restrict__ cell_of, int n_num, int n,
                       double *__restrict__ packed)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_num) return;
    int c = cell_of[r];
#pragma unroll
    for (int v = 0; v < 3; v++) packed[(size_t)v * n_num + r] = (double)planes[(size_t)v * n + c];
}

template <typename T>
__global__ void k_unpack(const double *__restrict__ packed, const int32_t *__restrict__ cell_of, int n_num, int n,
                         T *__restrict__ planes)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_num) return;
    int c = cell_of[r];
#pragma unroll
    for (int v = 0; v < 3; v++) planes[(size_t)v * n + c] = (T)packed[(size_t)v * n_num + r];
}

template <typename T>
void pack_launch(const T *planes, const int32_t *cell_of, int n_num, int n, double *packed, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pack<T>, dim3((n_num + 255) / 256), dim3(256), 0, stream, planes, cell_of, n_num, n, packed);
}

template <typename T>
void unpack_launch(const double *packed, const int32_t *cell_of, int n_num, int n, T *planes, hipStream_t stream)
{
    hipLaunchKernelGGL(k_unpack<T>, dim3((n_num + 255) / 256), dim3(256), 0, stream, packed, cell_of, n_num, n,
                       planes);
}

template <typename T>
__global__ void k_stimulate(T *__restrict__ planes, const uint8_t *__restrict__ mask, int n, T value)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && mask[k]) planes[k] = value;
}

template <typename T>
void stimulate_launch(T *planes, const uint8_t *mask, int n, T value, hipStream_t stream)
{
    hipLaunchKernelGGL(k_stimulate<T>, dim3((n + 255) / 256), dim3(256), 0, stream, planes, mask, n, value);
}

template void objective_launch<double>(int, const int64_t *, const int32_t *, const double *, double, const int32_t *,
                                       int, const double *, const double *, double *, double *, hipStream_t);
template void objective_launch<float>(int, const int64_t *, const int32_t *, const double *, double, const int32_t *,
                                      int, const float *, const double *, float *, double *, hipStream_t);
template void pack_launch<double>(const double *, const int32_t *, int, int, double *, hipStream_t);
template void pack_launch<float>(const float *, const int32_t *, int, int, double *, hipStream_t);
template void unpack_launch<double>(const double *, const int32_t *, int, int, double *, hipStream_t);
template void unpack_launch<float>(const double *, const int32_t *, int, int, float *, hipStream_t);
template void stimulate_launch<double>(double *, const uint8_t *, int, double, hipStream_t);
template void stimulate_launch<float>(float *, const uint8_t *, int, float, hipStream_t);

}  // namespace fk

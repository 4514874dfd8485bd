// Vector and matrix-vector kernels of the mixed-precision solver (iterative refinement and
// GMRES-IR): precision conversions, the FP64 residual / product with A, norms and inner
// products. Every reduction has a fixed order (one workgroup per result), so results are
// reproducible run to run.
#pragma once

#include "lrn_device.hpp"

namespace lrn {

template <typename Tin, typename Tout>
__global__ void convert_kernel(const Tin* __restrict__ in, Tout* __restrict__ out, long long count)
{
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        out[i] = static_cast<Tout>(in[i]);
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// out[i] = beta * b[i] + alpha * (A x)[i] (b may be null: treated as 0), one wave per row,
// FP64 with fused multiply-adds. mode 1 instead stores sum_j |A[i][j]| (for ||A||_inf).
__global__ __launch_bounds__(256) void gemv_rows_kernel(const double* __restrict__ A, long long n,
    const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ out, double alpha, double beta,
    int mode)
{
    const int lane = threadIdx.x & 63;
    const long long row = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* a = A + row * n;
    double s = 0.0;
    if (mode == 0) {
        for (long long c = lane; c < n; c += 64) s = fma(a[c], x[c], s);
    } else {
        for (long long c = lane; c < n; c += 64) s += fabs(a[c]);
    }
    s = wave_sum_f64(s);
    if (lane == 0) {
        if (mode == 0)
            out[row] = (b != nullptr ? beta * b[row] : 0.0) + alpha * s;
        else
            out[row] = s;
    }
}

// The FP32 copy of A (n x n, row-major, as convert_kernel) fused with the row sums of |A| for
// ||A||_inf, so A is read once for both: one wave per row. kVec (even n, 16-byte aligned A):
// 16-byte loads and 8-byte stores, lane l taking the column pairs 2 l + 128 k; otherwise one
// element per lane and step. Each row's sum has a fixed order, so the norm is reproducible.
template <bool kVec>
__global__ __launch_bounds__(256) void convert_rowsum_kernel(const double* __restrict__ A, long long n,
    float* __restrict__ out, double* __restrict__ rowsum)
{
    const int lane = threadIdx.x & 63;
    const long long row = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* a = A + row * n;
    float* o = out + row * n;
    double s = 0.0;
    if (kVec) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        typedef float f2 __attribute__((ext_vector_type(2)));
#pragma unroll 4
        for (long long c = 2 * lane; c < n; c += 128) {
            const d2 v = *reinterpret_cast<const d2*>(a + c);
            f2 w;
            w.x = static_cast<float>(v.x);
            w.y = static_cast<float>(v.y);
            *reinterpret_cast<f2*>(o + c) = w;
            s += fabs(v.x) + fabs(v.y);
        }
    } else {
#pragma unroll 4
        for (long long c = lane; c < n; c += 64) {
            const double v = a[c];
            o[c] = static_cast<float>(v);
            s += fabs(v);
        }
    }
    s = wave_sum_f64(s);
    if (lane == 0) rowsum[row] = s;
}

// The same product (mode 0) with 16-byte loads, four in flight per lane and four partial sums
// per lane (even n, 16-byte aligned A and x): 0.35 ms against 0.41 ms at n = 16384 (6.1 TB/s).
typedef double gemv_d2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void gemv_rows_vec_kernel(const double* __restrict__ A, long long n,
    const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ out, double alpha, double beta)
{
    const int lane = threadIdx.x & 63;
    const long long row = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const gemv_d2* a = reinterpret_cast<const gemv_d2*>(A + row * n);
    const gemv_d2* xv = reinterpret_cast<const gemv_d2*>(x);
    const long long m = n / 2;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    long long c = lane;
    for (; c + 192 < m; c += 256) {
        const gemv_d2 a0 = a[c], a1 = a[c + 64], a2 = a[c + 128], a3 = a[c + 192];
        const gemv_d2 x0 = xv[c], x1 = xv[c + 64], x2 = xv[c + 128], x3 = xv[c + 192];
        s0 = fma(a0.x, x0.x, s0);
        s0 = fma(a0.y, x0.y, s0);
        s1 = fma(a1.x, x1.x, s1);
        s1 = fma(a1.y, x1.y, s1);
        s2 = fma(a2.x, x2.x, s2);
        s2 = fma(a2.y, x2.y, s2);
        s3 = fma(a3.x, x3.x, s3);
        s3 = fma(a3.y, x3.y, s3);
    }
    for (; c < m; c += 64) {
        const gemv_d2 a0 = a[c], x0 = xv[c];
        s0 = fma(a0.x, x0.x, s0);
        s0 = fma(a0.y, x0.y, s0);
    }
    double s = (s0 + s1) + (s2 + s3);
    s = wave_sum_f64(s);
    if (lane == 0) out[row] = (b != nullptr ? beta * b[row] : 0.0) + alpha * s;
}

// Single-workgroup reductions of length-n vectors (deterministic order).
// op 0: max |u|; op 1: sum u*v; op 2: sum u*u.
__global__ __launch_bounds__(1024) void reduce_kernel(const double* __restrict__ u, const double* __restrict__ v,
    long long n, int op, double* __restrict__ out)
{
    __shared__ double part[16];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) {
        const double a = u[i];
        if (op == 0)
            s = fmax(s, fabs(a));
        else if (op == 1)
            s = fma(a, v[i], s);
        else
            s = fma(a, a, s);
    }
    if (op == 0) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
    } else {
        s = wave_sum_f64(s);
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = part[0];
        for (int q = 1; q < 16; ++q) t = (op == 0) ? fmax(t, part[q]) : t + part[q];
        *out = t;
    }
}

// h[k] = sum_i V[k][i] * w[i] for k < nv (one workgroup per k); V holds vectors of length n
// with stride ldv.
__global__ __launch_bounds__(256) void multi_dot_kernel(const double* __restrict__ V, long long ldv,
    const double* __restrict__ w, long long n, double* __restrict__ h)
{
    __shared__ double part[4];
    const double* v = V + static_cast<long long>(blockIdx.x) * ldv;
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) s = fma(v[i], w[i], s);
    s = wave_sum_f64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) h[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// w[i] -= sum_k V[k][i] * h[k], k < nv
__global__ void multi_axpy_kernel(const double* __restrict__ V, long long ldv, const double* __restrict__ h, int nv,
    double* __restrict__ w, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int k = 0; k < nv; ++k) s = fma(-V[static_cast<long long>(k) * ldv + i], h[k], s);
    w[i] = s;
}

// y = alpha * y + beta * x (alpha / beta from host values, or from device scalars when the
// pointers are non-null: y = y * (*pa) + x * (*pb))
__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, long long n, double alpha,
    double beta, const double* __restrict__ pa, const double* __restrict__ pb)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = pa ? *pa : alpha;
    const double b = pb ? *pb : beta;
    y[i] = a * y[i] + (x ? b * x[i] : 0.0);
}

// y += (double) d  (FP32 correction added to the FP64 iterate)
__global__ void add_f32_kernel(double* __restrict__ y, const float* __restrict__ d, long long n)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += static_cast<double>(d[i]);
}

// v = w / s with s a device scalar (GMRES normalisation)
__global__ void scale_inv_kernel(double* __restrict__ v, const double* __restrict__ w, long long n,
    const double* __restrict__ s)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = w[i] / *s;
}

}  // namespace lrn

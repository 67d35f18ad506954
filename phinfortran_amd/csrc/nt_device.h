// nt_device.h -- device-side building blocks shared by the transport kernels:
// interleaved complex storage, batched matrix addressing and the CDNA4 MFMA
// wrappers used for complex products.
//
// Storage convention everywhere: complex matrices are column-major with a
// leading dimension in complex elements, re/im interleaved (the layout of the
// reference's complex(kind=8) arrays and of numpy complex128 in F order).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nt {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct vec2;
template <> struct vec2<double> { typedef double2 type; };
template <> struct vec2<float> { typedef float2 type; };

template <typename T> struct acc4;
template <> struct acc4<double> { typedef v4d type; };
template <> struct acc4<float> { typedef v4f type; };

// D(16x16) += A(16x4) * B(4x16).  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; it receives D[mfma_out_row(l, r)][l%16] in element r.  The
// row of element r differs between the FP64 and FP32 forms: the FP64 output is
// four row groups of one register (row 4r + l/16), the FP32 output one group
// of four consecutive rows (row 4(l/16) + r).
__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T> __device__ __forceinline__ int mfma_out_row(int lane, int r);
template <> __device__ __forceinline__ int mfma_out_row<double>(int lane, int r) { return 4 * r + (lane >> 4); }
template <> __device__ __forceinline__ int mfma_out_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

// One batch of matrices.  Item p (0 <= p < batch) of group g lives at
//   base + ((idx ? idx[p] : p) * stride + g * gstride) complex elements.
// stride 0 shares one matrix across the batch (e.g. H(k) blocks across
// energies); idx maps points to k-points.  idx_n (host-side only) is the
// number of distinct index values, used by the launch-time bounds audit.
template <typename T> struct BMat {
    const T* base;
    long stride;
    long gstride;
    const int* idx;
    int ld;
    int idx_n;
};
template <typename T> struct BMatW {
    T* base;
    long stride;
    long gstride;
    const int* idx;
    int ld;
    int idx_n;
};

template <typename T>
__device__ __forceinline__ const T* bptr(const BMat<T>& m, int p, int g)
{
    long i = m.idx ? (long)m.idx[p] : (long)p;
    return m.base + 2 * (i * m.stride + (long)g * m.gstride);
}
template <typename T>
__device__ __forceinline__ T* bptr(const BMatW<T>& m, int p, int g)
{
    long i = m.idx ? (long)m.idx[p] : (long)p;
    return m.base + 2 * (i * m.stride + (long)g * m.gstride);
}

// Lanes of one wave64.
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

}  // namespace nt

// Shared device helpers for the rlm/hlm HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define RLM_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            fprintf(stderr, "[rlm_hip] %s failed: %s (%s:%d)\n", #expr,            \
                    hipGetErrorString(_e), __FILE__, __LINE__);                    \
            return -(int)_e;                                                       \
        }                                                                          \
    } while (0)

namespace rlm {

constexpr int kWave = 64;

// ---------------------------------------------------------------- complex (double2)
struct cd {
    double x, y;
};

__host__ __device__ inline cd mk(double a, double b) { return cd{a, b}; }
__host__ __device__ inline cd cadd(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd csub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd cmul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// conj(a) * b
__host__ __device__ inline cd cmulc(cd a, cd b) { return cd{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }
__host__ __device__ inline cd cconj(cd a) { return cd{a.x, -a.y}; }
__host__ __device__ inline cd cscale(cd a, double s) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline double cabs2(cd a) { return a.x * a.x + a.y * a.y; }
// acc += a*b
__host__ __device__ inline void cfma(cd &acc, cd a, cd b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a)*b
__host__ __device__ inline void cfmac(cd &acc, cd a, cd b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}
// 1/a
__host__ __device__ inline cd cinv(cd a) {
    double s = 1.0 / (a.x * a.x + a.y * a.y);
    return cd{a.x * s, -a.y * s};
}

__device__ inline cd ld(const double *p, long i) {
    const double2 v = reinterpret_cast<const double2 *>(p)[i];
    return cd{v.x, v.y};
}
__device__ inline void st(double *p, long i, cd v) {
    reinterpret_cast<double2 *>(p)[i] = make_double2(v.x, v.y);
}

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum of one double; `red` is LDS scratch of >= blockDim/64 doubles.
// All threads receive the total. Contains two barriers.
__device__ inline double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; i++) t += red[i];
    __syncthreads();
    return t;
}

// Block-wide sum of a complex value.
__device__ inline cd block_sum_c(cd v, double *red) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) {
        red[2 * wid] = v.x;
        red[2 * wid + 1] = v.y;
    }
    __syncthreads();
    cd t{0.0, 0.0};
    for (int i = 0; i < nw; i++) {
        t.x += red[2 * i];
        t.y += red[2 * i + 1];
    }
    __syncthreads();
    return t;
}

}  // namespace rlm

// ---------------------------------------------------------------- precision-generic complex
namespace rlm {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <typename T>
__host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>{a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
    return cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ inline cx<T> operator*(cx<T> a, T s) { return cx<T>{a.x * s, a.y * s}; }
template <typename T>
__host__ __device__ inline cx<T> conjx(cx<T> a) { return cx<T>{a.x, -a.y}; }
template <typename T>
__host__ __device__ inline T abs2x(cx<T> a) { return a.x * a.x + a.y * a.y; }
// acc += a*b
template <typename T>
__host__ __device__ inline void fmax_(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a)*b
template <typename T>
__host__ __device__ inline void fmacx(cx<T> &acc, cx<T> a, cx<T> b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}
template <typename T>
__device__ inline cx<T> ldx(const T *p, long i) { return reinterpret_cast<const cx<T> *>(p)[i]; }

// Visits (a, b), a < na, b < nb, in the order of the flat index a * nb + b = lane + k G
// (the same element per lane as that loop), carrying (a, b) from one element to the next
// instead of dividing the flat index: two divisions per lane instead of one per element.
template <typename F>
__device__ __forceinline__ void for_2d(int lane, int G, int na, int nb, F f) {
    if (na <= 0 || nb <= 0) return;
    int a = lane / nb, b = lane - a * nb;
    const int da = G / nb, db = G - da * nb;
    while (a < na) {
        f(a, b);
        a += da;
        b += db;
        if (b >= nb) {
            b -= nb;
            a++;
        }
    }
}

template <typename T>
__device__ inline void stx(T *p, long i, cx<T> v) { reinterpret_cast<cx<T> *>(p)[i] = v; }

template <typename T>
__device__ inline T wave_sum_t(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum; `red` is LDS scratch of >= 2*blockDim/64 elements of T. Two barriers.
template <typename T>
__device__ inline T block_sum_t(T v, T *red) {
    v = wave_sum_t(v);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    T t = 0;
    for (int i = 0; i < nw; i++) t += red[i];
    __syncthreads();
    return t;
}

template <typename T>
__device__ inline cx<T> block_sum_cx(cx<T> v, T *red) {
    v.x = wave_sum_t(v.x);
    v.y = wave_sum_t(v.y);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (lane == 0) {
        red[2 * wid] = v.x;
        red[2 * wid + 1] = v.y;
    }
    __syncthreads();
    cx<T> t{0, 0};
    for (int i = 0; i < nw; i++) {
        t.x += red[2 * i];
        t.y += red[2 * i + 1];
    }
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------- matrix cores
// Real 16x16x4 MFMAs (four per complex product). Lane l of an MFMA holds
// A[l%16][l/16] and B[l/16][l%16]; the 16x16 result D sits in 4 elements per lane, element v of lane l being
// D[Mf<T>::row(l, v)][l%16]. Element v of an accumulator is therefore directly
// the B operand of a K-step whose K index l/16 stands for row Mf<T>::row(l, v),
// so chained products need no data movement between them.
template <typename T>
struct Mf;
template <>
struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int l, int v) { return (l >> 4) + 4 * v; }
};
template <>
struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ inline int row(int l, int v) { return 4 * (l >> 4) + v; }
};


// ---------------------------------------------------------------- band form
// Circular band D = |c - r| in B = {0..o} U {N-o..N-1} (all of [0, N) when
// 2o+1 >= N), stored as b = 0..nb-1: band_D maps b to D, band_b maps D to b (-1
// outside the band).
__host__ __device__ inline int band_nb(int N, int off) { return 2 * off + 1 >= N ? N : 2 * off + 1; }
__device__ inline int band_D(int b, int N, int off) { return (2 * off + 1 >= N || b <= off) ? b : b + N - 2 * off - 1; }
__device__ inline int band_b(int D, int N, int off) {
    if (2 * off + 1 >= N || D <= off) return D;
    return D >= N - off ? D - N + 2 * off + 1 : -1;
}

// Entry (r, c) of a circular-band matrix: from the dense column-major matrix Ak when
// Hk is null, else from the band form Hk (n x nb, offset bw) with the convention of
// the dense expansion: conj(H[r][b(c - r)]) for c >= r, H[c][b(r - c)] below.
__device__ inline cd band_entry(const double *Ak, const double *Hk, int n, int bw, int nb, int r, int c) {
    if (!Hk) return ld(Ak, (long)c * n + r);
    if (r <= c) {
        const int b = band_b(c - r, n, bw);
        return b >= 0 ? cconj(ld(Hk, (long)r * nb + b)) : cd{0.0, 0.0};
    }
    const int b = band_b(r - c, n, bw);
    return b >= 0 ? ld(Hk, (long)c * nb + b) : cd{0.0, 0.0};
}

}  // namespace rlm

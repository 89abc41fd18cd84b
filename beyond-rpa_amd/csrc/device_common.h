// Device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef short s8_t __attribute__((ext_vector_type(8)));

// Fragment index f with pre[f] <= b < pre[f+1] (pre has nfrag+1 entries).
__device__ __forceinline__ int find_frag(const int64_t *__restrict__ pre, int nfrag, int64_t b) {
    int lo = 0, hi = nfrag - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The same search on a copy of pre staged in LDS (one parallel load instead of
// a chain of dependent global loads; matters for small work-groups).  Every
// thread of the work-group must call it (it synchronises).
constexpr int FIND_LDS_MAX = 512;
__device__ __forceinline__ int find_frag_lds(const int64_t *__restrict__ pre, int nfrag, int64_t b,
                                             int64_t *s_pre) {
    if (nfrag > FIND_LDS_MAX) return find_frag(pre, nfrag, b);
    for (int i = threadIdx.x; i <= nfrag; i += blockDim.x) s_pre[i] = pre[i];
    __syncthreads();
    return find_frag(s_pre, nfrag, b);
}

// Lower-triangle pair index -> (i, j), i >= j, pair = i*(i+1)/2 + j.
__device__ __forceinline__ void tri_decode(int pair, int &i, int &j) {
    int t = (int)((sqrt(8.0 * (double)pair + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= pair) ++t;
    while (t * (t + 1) / 2 > pair) --t;
    i = t;
    j = pair - t * (t + 1) / 2;
}

// Double-double accumulation (Knuth TwoSum), used where the reference sums
// in REAL(F128).
struct dd_t {
    double hi, lo;
};
__device__ __forceinline__ dd_t dd_add(dd_t a, double b) {
    double s = a.hi + b;
    double bb = s - a.hi;
    double err = (a.hi - (s - bb)) + (b - bb);
    double lo = a.lo + err;
    double hi = s + lo;
    lo = lo - (hi - s);
    return {hi, lo};
}
__device__ __forceinline__ dd_t dd_add(dd_t a, dd_t b) {
    dd_t s = dd_add(a, b.hi);
    return dd_add(s, b.lo);
}

// XCD-aware block order.  Work-group ids are handed to the 8 XCDs round-robin
// (id % 8), and each XCD has its own L2.  This bijection on [0, nblocks) gives
// XCD x a contiguous range of logical work items, so tiles that share panels
// (neighbouring items of one matrix) run on the same XCD and hit its L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nblocks) {
    constexpr int64_t NX = 8;
    const int64_t q = nblocks / NX, rem = nblocks % NX;
    const int64_t x = b % NX, r = b / NX;
    return x * q + (x < rem ? x : rem) + r;
}

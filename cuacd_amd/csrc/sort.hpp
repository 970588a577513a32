// sort.hpp - workgroup-cooperative sorts in LDS.
//
// Bitonic sorting network over a power-of-two LDS buffer: every stage is one
// pass of P/2 independent compare-exchanges spread over the workgroup, then a
// barrier. Used for small and medium arrays (up to a few thousand records);
// padding slots hold a caller-supplied sentinel that sorts last.
#pragma once
#include "wave.hpp"

namespace cuacd {

__host__ __device__ constexpr int pow2_ceil(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Sorts lds[0..P) ascending by `less`. All threads of the block must call.
template <class T, class Less>
__device__ void block_bitonic_sort(T* lds, int P, Less less) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < (P >> 1); i += nt) {
                // i-th compare-exchange of this stage: lower index a, partner b = a + j
                const int a = ((i / j) * (j << 1)) + (i % j);
                const int b = a + j;
                const bool up = (a & k) == 0;
                T x = lds[a], y = lds[b];
                if (less(y, x) == up) {
                    lds[a] = y;
                    lds[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace cuacd

// Device core of the batched FFT, shared by the plain transforms (g3d_fft.hip) and the fused kernels of the plasma
// pass (g3d_pass.hip): bit-reversed load into a padded LDS layout, then the radix-2 decimation-in-time stages three
// at a time in registers.
#pragma once
#include "g3d_common.hpp"

namespace g3d {
namespace fftcore {

__device__ __forceinline__ int bitrev(int i, int logn) { return logn ? int(__brev(unsigned(i)) >> (32 - logn)) : 0; }

// LDS position of element i of a sequence: one complex of padding per 8, so that the strided group accesses
// below fall into different banks
__device__ __forceinline__ int lpos(int i) { return i + (i >> 3); }

// K radix-2 DIT stages s0 .. s0+K-1 for one closed group of 2^K elements held in registers: the members
// h 2^(s0-1+K) + m 2^(s0-1) + r, m = 0 .. 2^K-1.  Same butterflies, same twiddles, same order of operations as
// the reference's iterative transform (src/fft.hpp:30-33, 74-85), so the result is bit for bit the same.
template <int K, class C>
__device__ __forceinline__ void group_stages(C* seq, int n, int s0, int h, int r, const C* tws) {
    constexpr int G = 1 << K;
    const int low = s0 - 1;
    const int base = (h << (low + K)) + r;
    C v[G];
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = seq[lpos(base + (m << low))];
#pragma unroll
    for (int t = 0; t < K; ++t) {
        const int half = 1 << t, step = n >> (s0 + t);
#pragma unroll
        for (int m = 0; m < G; ++m) {
            if (m & half) continue;
            const int j = ((m & (half - 1)) << low) + r;
            const C w = tws[j * step];
            const C u = v[m], x = v[m + half] * w;
            v[m] = u + x;
            v[m + half] = u - x;
        }
    }
#pragma unroll
    for (int m = 0; m < G; ++m) seq[lpos(base + (m << low))] = v[m];
}

// the transform of nseq sequences of length n = 2^L in LDS (bit-reversed order, padded layout, sequence q at
// q * spad): the stages in groups of three, one LDS round trip and one barrier per group
template <class C>
__device__ __forceinline__ void lds_fft(C* buf, int n, int L, int nseq, int spad, const C* tws) {
    for (int s0 = 1; s0 <= L; s0 += 3) {
        const int K = L - s0 + 1 < 3 ? L - s0 + 1 : 3;
        const int lgn = L - K;                       // log2 of the groups per sequence
        const int low = s0 - 1;
        const int tot = nseq << lgn;
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            const int q = t >> lgn, g = t & ((1 << lgn) - 1);
            const int r = g & ((1 << low) - 1), h = g >> low;
            C* seq = buf + q * spad;
            if (K == 3) group_stages<3, C>(seq, n, s0, h, r, tws);
            else if (K == 2) group_stages<2, C>(seq, n, s0, h, r, tws);
            else group_stages<1, C>(seq, n, s0, h, r, tws);
        }
        __syncthreads();
    }
}

// the double twiddle table in the transform's precision (conjugated for the inverse)
template <class C>
__device__ __forceinline__ void load_twiddles(C* tws, const cx* tw, int hn, bool inv) {
    for (int k = threadIdx.x; k < hn; k += blockDim.x) {
        cx w = tw[k];
        if (inv) w.im = -w.im;
        tws[k] = twiddle_as<C>(w);
    }
}

__device__ __forceinline__ int padded(int n) { return n + (n >> 3); }

// The row block of workgroup b of nb when the row blocks of one XCD are to be consecutive: workgroup b runs on XCD
// b mod 8, so block (b mod 8) nb/8 + b/8 puts the blocks [x nb/8, (x + 1) nb/8) on XCD x, and the transposed
// 8-byte accesses of neighbouring rows (mode ix ny + j) meet in that XCD's L2 as whole lines instead of partial
// ones (measured at 256 x 256: the rows-forward kernel writes 2.9 MB instead of 7.6 MB)
__device__ __forceinline__ int row_block(int b, int nb, int xcd) {
    return (xcd && (nb & 7) == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
}

}  // namespace fftcore
}  // namespace g3d

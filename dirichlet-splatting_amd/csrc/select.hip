// Certificate scores of a batched scan and the top candidates of every scan, on the device
// (reference: dsplat/certificate.py:11-31 for the score; the ranking is the stable descending
// argsort that certificate.py:52-56 and :196-201 apply to scores).
#include <limits.h>
#include <stdint.h>

#include "dsplat_internal.h"

namespace dsplat {

constexpr int kTopThreads = 1024;
constexpr int kTopWaves = kTopThreads / 64;

// score = |s|^2 area / max(norm area, 1e-60) / power, s = corr or Re corr, in float64 from
// the float32 (corr.re, corr.im, norm) rows of the scan. The floor keeps NaN, as a clamp does.
__global__ __launch_bounds__(256) void certificate_score_kernel(const float* __restrict__ scan, int C,
                                                                const float* __restrict__ areas,
                                                                const double* __restrict__ power, int real_part,
                                                                double* __restrict__ scores) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= C) return;
    const float* v = scan + ((size_t)n * C + i) * 3;
    const double re = v[0], im = real_part ? 0.0 : (double)v[1];
    const double area = areas[i];
    const double s2 = (re * re + im * im) * area;
    double norm = (double)v[2] * area;
    norm = norm < 1e-60 ? 1e-60 : norm;
    scores[(size_t)n * C + i] = s2 / norm / power[n];
}

// An unsigned key with the order of the doubles: NaN above everything (first in a descending
// sort), -0 equal to +0.
__device__ __forceinline__ uint64_t order_key(double x) {
    if (x != x) return ~0ull;
    const uint64_t b = x == 0.0 ? 0ull : (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

// before(a, b): a ranks first, by key descending and index ascending.
__device__ __forceinline__ bool ranks_before(uint64_t ka, int ia, uint64_t kb, int ib) {
    return ka > kb || (ka == kb && ia < ib);
}

// One workgroup per row: the k = min(top, count) largest keys by an 8-bit radix select from
// the top byte down (each pass histograms the keys that share the digits chosen so far), then
// every key above the threshold T and the lowest-index ones equal to it are gathered into LDS
// and bitonic-sorted by (key desc, index asc).
__global__ __launch_bounds__(kTopThreads) void top_descending_kernel(const double* __restrict__ scores, int count,
                                                                     int k, int64_t* __restrict__ out) {
    __shared__ uint64_t keys[kTopMax];
    __shared__ int ids[kTopMax];
    __shared__ int hist[256];
    __shared__ int wave_count[kTopWaves];
    __shared__ int state[4];  // chosen digit, still to take, equal to the threshold, gather slot
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* x = scores + (size_t)blockIdx.x * count;

    uint64_t prefix = 0, mask = 0;
    int remaining = k, equal = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        for (int b = 0; b < count; b += kTopThreads) {
            const int i = b + tid;
            const uint64_t key = i < count ? order_key(x[i]) : 0;
            const bool match = i < count && (key & mask) == prefix;
            const int bin = (int)((key >> shift) & 255);
            // one LDS atomic per wave when every matching lane falls in one bin (the common
            // case of the first passes, where most keys share their top digits)
            const uint64_t active = __ballot(match);
            if (active) {
                const int lead = __ffsll((unsigned long long)active) - 1;
                const int first = __shfl(bin, lead, 64);
                if (__ballot(match && bin == first) == active) {
                    if (lane == lead) atomicAdd(&hist[first], __popcll(active));
                } else if (match) {
                    atomicAdd(&hist[bin], 1);
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            // the digit d where the keys above it number fewer than `remaining` and those at
            // or above it at least as many; the keys matching the prefix number >= remaining
            int above = 0, d = 255;
            for (; d > 0; --d) {
                if (above + hist[d] >= remaining) break;
                above += hist[d];
            }
            state[0] = d;
            state[1] = remaining - above;
            state[2] = hist[d];
        }
        __syncthreads();
        prefix |= (uint64_t)state[0] << shift;
        mask |= 255ull << shift;
        remaining = state[1];
        equal = state[2];
        __syncthreads();
    }
    const uint64_t T = prefix;
    const int need = remaining;  // keys equal to T to take: 1 <= need <= equal
    const bool all_equal = need == equal;
    if (tid == 0) state[3] = 0;
    __syncthreads();
    for (int i = tid; i < count; i += kTopThreads) {
        const uint64_t key = order_key(x[i]);
        if (key > T || (all_equal && key == T)) {
            const int slot = atomicAdd(&state[3], 1);
            keys[slot] = key;
            ids[slot] = i;
        }
    }
    __syncthreads();
    if (!all_equal) {
        // ties at T: the `need` lowest indices, by a workgroup prefix count in index order
        const int base = k - need;
        int taken = 0;
        for (int b = 0; b < count && taken < need; b += kTopThreads) {
            const int i = b + tid;
            const bool eq = i < count && order_key(x[i]) == T;
            const uint64_t ball = __ballot(eq);
            if (lane == 0) wave_count[wave] = __popcll(ball);
            __syncthreads();
            int before = 0, total = 0;
            for (int w = 0; w < kTopWaves; ++w) {
                const int c = wave_count[w];
                before += w < wave ? c : 0;
                total += c;
            }
            const int rank = taken + before + __popcll(ball & ((1ull << lane) - 1));
            if (eq && rank < need) {
                keys[base + rank] = T;
                ids[base + rank] = i;
            }
            taken += total;
            __syncthreads();
        }
    }
    int P = 1;
    while (P < k) P <<= 1;
    for (int i = k + tid; i < P; i += kTopThreads) {
        keys[i] = 0;  // below every real key (NaN maps to the top, no double maps to 0)
        ids[i] = INT_MAX;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < P; i += kTopThreads) {
                const int j = i ^ stride;
                if (j > i) {
                    const uint64_t ki = keys[i], kj = keys[j];
                    const int ii = ids[i], ij = ids[j];
                    const bool first_half = (i & size) == 0;  // ranks ascending by `before` here
                    const bool swap = first_half ? ranks_before(kj, ij, ki, ii) : ranks_before(ki, ii, kj, ij);
                    if (swap) {
                        keys[i] = kj;
                        keys[j] = ki;
                        ids[i] = ij;
                        ids[j] = ii;
                    }
                }
            }
            __syncthreads();
        }
    for (int r = tid; r < k; r += kTopThreads) out[(size_t)blockIdx.x * k + r] = ids[r];
}

hipError_t launch_certificate_scores(const float* scan, int n, int C, const float* areas, const double* power,
                                     int real_part, double* scores, hipStream_t st) {
    const dim3 grid((C + 255) / 256, n);
    certificate_score_kernel<<<grid, 256, 0, st>>>(scan, C, areas, power, real_part, scores);
    return hipGetLastError();
}

hipError_t launch_top_descending(const double* scores, int rows, int count, int top, int64_t* out, hipStream_t st) {
    const int k = top < count ? top : count;
    top_descending_kernel<<<rows, kTopThreads, 0, st>>>(scores, count, k, out);
    return hipGetLastError();
}

}  // namespace dsplat

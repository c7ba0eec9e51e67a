// Candidate kernels of the spatial search: range gate, greedy separated peaks and the
// batched quasi-Newton direction and trial step (reference: dsplat/kernels/search.slang:1-109,
// bindings dsplat/kernels/search.py:1-77).
#include <float.h>
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "dsplat_internal.h"

namespace dsplat {

// A point passes when its spherical range bin (round half to even) falls in a marked
// interval at `required` apertures or more (search.slang:7-24).
__global__ __launch_bounds__(256) void range_gate_kernel(const float* __restrict__ pts, const float* __restrict__ pos,
                                                         const int* __restrict__ intervals, int count, int apertures,
                                                         int bins, float spacing, int required,
                                                         int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    int matches = 0;
    for (int a = 0; a < apertures; ++a) {
        const float dx = px - pos[3 * a], dy = py - pos[3 * a + 1], dz = pz - pos[3 * a + 2];
        const float x = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz))) / spacing;
        const float lo = floorf(x);
        int k = (int)lo;
        const float frac = x - lo;
        if (frac > 0.5f || (frac == 0.5f && (k & 1))) ++k;
        if (k >= 0 && k < bins && intervals[(size_t)a * bins + k] != 0) ++matches;
    }
    out[i] = matches >= required ? 1 : 0;
}

// Greedy best-first separated peaks (search.slang:27-64): repeatedly take the best finite
// score (ties to the lower index) and drop every candidate closer than the separation.
// The same selection is made here from the candidates sorted by (score desc, index asc):
// a candidate is taken iff no earlier taken one lies within the separation. Distances are
// formed exactly as in the serial definition, d = c_candidate - c_taken accumulated by fmaf
// over the coordinates.
//
// One workgroup. Windows of 2048 sorted candidates are first tested in parallel against the
// peaks taken so far, and the survivors go, in sorted order, into an LDS queue: a candidate
// close to a taken peak can never be taken, so this only skips work. Blocks of 256 queued
// survivors are then resolved at a time: each is tested again against every peak taken
// since, each records which later candidates of the block clash with it (bitmasks), and
// one wave walks the block in order, one turn per taken peak. Past the first peaks most
// sorted candidates sit next to a taken one, so most windows add few survivors and the
// blocks (several barriers each) stay few.

constexpr int kPeakBlock = 256;
constexpr int kPeakLdsFloats = 8192;  // taken centers cached in LDS (32 KB); the rest in the spill
constexpr int kPeakPer = 4;           // window candidates per thread
constexpr int kPeakChainLds = 4096;   // bucket chain links of the first taken peaks in LDS
constexpr int kPeakWindow = kPeakPer * kPeakBlock;
constexpr int kPeakQueue = kPeakWindow + kPeakBlock;  // fewer than 256 left over plus one window
constexpr int kPeakMaxDim = 8;

__global__ __launch_bounds__(256) void peak_keys(const float* __restrict__ scores, int count,
                                                 float* __restrict__ keys, int* __restrict__ ids) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float v = scores[i];
    keys[i] = isfinite(v) ? v + 0.0f : -INFINITY;  // +0.0 folds -0.0 into 0.0: equal scores tie
    ids[i] = i;
}

__device__ __forceinline__ float peak_d2(const float* a, const float* b, int dim) {
    float d2 = 0.0f;
#pragma unroll
    for (int j = 0; j < kPeakMaxDim; ++j)
        if (j < dim) {
            const float d = a[j] - b[j];
            d2 = fmaf(d, d, d2);
        }
    return d2;
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int lane) {
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, lane);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// Uniform grid over the first (up to 3) coordinates for the taken peaks: cells of edge
// 2.02 separation. A peak closer than the separation to point p differs from it by less than
// 0.495 cell edges per coordinate, so it lies in p's cell or in the neighbour on the side of
// p's nearer cell face: 2 cells per coordinate (8 in three dimensions) even after rounding;
// buckets hold chains of taken indices (hash collisions are weeded out by the distance test
// itself).
constexpr int kPeakBuckets = 4096;

__device__ __forceinline__ int peak_bucket(int ix, int iy, int iz) {
    return (int)(((unsigned)ix * 73856093u ^ (unsigned)iy * 19349663u ^ (unsigned)iz * 83492791u) &
                 (kPeakBuckets - 1));
}

__global__ __launch_bounds__(kPeakBlock) void peaks_kernel(const float* __restrict__ keys,
                                                           const int* __restrict__ order,
                                                           const float* __restrict__ ctr, int count, int peaks,
                                                           int dim, float sep2, float* __restrict__ spill,
                                                           int* __restrict__ chain, int* __restrict__ out) {
    __shared__ float taken_lds[kPeakLdsFloats];
    __shared__ int head[kPeakBuckets];
    __shared__ int chain_lds[kPeakChainLds];  // the rest in `chain`
    __shared__ float block_c[kPeakBlock * kPeakMaxDim];
    __shared__ unsigned long long conflict[kPeakBlock][kPeakBlock / 64];
    __shared__ unsigned long long chosen[kPeakBlock / 64];
    __shared__ int alive[kPeakBlock];
    __shared__ int q_id[kPeakQueue];                   // queued survivors: candidate index
    __shared__ float q_c[kPeakQueue * kPeakMaxDim];    // and its coordinates
    __shared__ int offsets[kPeakPer * (kPeakBlock / 64) + 1];
    __shared__ int n_taken, finished, window_end;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lds_cap = kPeakLdsFloats / dim;
    const bool grid = dim <= 3;  // taken lookups through the cell grid, else a linear scan
    const float inv_h = 1.0f / (2.02f * sqrtf(sep2));
    if (tid == 0) {
        n_taken = 0;
        finished = 0;
    }
    for (int i = tid; i < kPeakBuckets; i += kPeakBlock) head[i] = -1;
    auto cell = [&](const float* p, int j) {  // clamped far from int overflow
        return j < dim ? (int)fminf(fmaxf(floorf(p[j] * inv_h), -1e9f), 1e9f) : 0;
    };
    // whether c lies within the separation of a peak taken so far (n of them)
    auto clashes = [&](const float* c, int n) {
        if (!(sep2 > 0.0f)) return false;
        if (grid) {
            const int ix = cell(c, 0), iy = cell(c, 1), iz = cell(c, 2);
            // the second cell per coordinate: toward the nearer face (none past the dimension)
            auto side = [&](int j, int i0) {
                return j < dim ? (c[j] * inv_h - (float)i0 < 0.5f ? -1 : 1) : 0;
            };
            const int sx = side(0, ix), sy = side(1, iy), sz = side(2, iz);
            for (int bz = 0; bz <= (sz != 0); ++bz)
                for (int by = 0; by <= (sy != 0); ++by)
                    for (int bx = 0; bx <= (sx != 0); ++bx)
                        for (int q = head[peak_bucket(ix + bx * sx, iy + by * sy, iz + bz * sz)]; q >= 0;
                             q = q < kPeakChainLds ? chain_lds[q] : chain[q]) {
                            const float* t = q < lds_cap ? taken_lds + q * dim : spill + (size_t)(q - lds_cap) * dim;
                            if (peak_d2(c, t, dim) < sep2) return true;
                        }
            return false;
        }
        for (int q = 0; q < n; ++q) {
            const float* t = q < lds_cap ? taken_lds + q * dim : spill + (size_t)(q - lds_cap) * dim;
            if (peak_d2(c, t, dim) < sep2) return true;
        }
        return false;
    };
    __syncthreads();
    int scan = 0, q_head = 0, q_len = 0;  // uniform
    bool scan_done = count <= 0;
    for (;;) {
        // refill the queue until it holds a block or the finite candidates are exhausted
        while (q_len - q_head < kPeakBlock && !scan_done) {
            const int left = q_len - q_head;
            int keep_id = -1;
            float keep_c[kPeakMaxDim];
            if (tid < left) {
                keep_id = q_id[q_head + tid];
#pragma unroll
                for (int j = 0; j < kPeakMaxDim; ++j)
                    if (j < dim) keep_c[j] = q_c[(q_head + tid) * dim + j];
            }
            const int n = n_taken;
            int w_id[kPeakPer];
            float w_c[kPeakPer][kPeakMaxDim];
            bool w_ok[kPeakPer], ended = false;
#pragma unroll
            for (int i = 0; i < kPeakPer; ++i) {
                const int p = scan + i * kPeakBlock + tid;
                const bool valid = p < count && keys[p] != -INFINITY;  // sorted: -inf (non-finite) last
                ended |= !valid;
                w_id[i] = valid ? order[p] : -1;
#pragma unroll
                for (int j = 0; j < kPeakMaxDim; ++j)
                    if (j < dim) w_c[i][j] = valid ? ctr[(size_t)w_id[i] * dim + j] : 0.0f;
                w_ok[i] = valid;
            }
#pragma unroll
            for (int i = 0; i < kPeakPer; ++i) w_ok[i] = w_ok[i] && !clashes(w_c[i], n);
            if (tid == 0) window_end = 0;
            __syncthreads();  // every thread has read the leftovers and n_taken
            if (ended) window_end = 1;
            if (tid < left) {
                q_id[tid] = keep_id;
#pragma unroll
                for (int j = 0; j < kPeakMaxDim; ++j)
                    if (j < dim) q_c[tid * dim + j] = keep_c[j];
            }
            unsigned long long bal[kPeakPer];
#pragma unroll
            for (int i = 0; i < kPeakPer; ++i) {
                bal[i] = __ballot(w_ok[i]);
                if (lane == 0) offsets[i * (kPeakBlock / 64) + wave] = __popcll(bal[i]);
            }
            __syncthreads();
            if (tid == 0) {  // exclusive scan in sorted order: chunk-major, wave-minor
                int sum = 0;
                for (int i = 0; i < kPeakPer * (kPeakBlock / 64); ++i) {
                    const int v = offsets[i];
                    offsets[i] = sum;
                    sum += v;
                }
                offsets[kPeakPer * (kPeakBlock / 64)] = sum;
            }
            __syncthreads();
            const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
            for (int i = 0; i < kPeakPer; ++i)
                if (w_ok[i]) {
                    const int e = left + offsets[i * (kPeakBlock / 64) + wave] + __popcll(bal[i] & below);
                    q_id[e] = w_id[i];
#pragma unroll
                    for (int j = 0; j < kPeakMaxDim; ++j)
                        if (j < dim) q_c[e * dim + j] = w_c[i][j];
                }
            q_head = 0;
            q_len = left + offsets[kPeakPer * (kPeakBlock / 64)];
            scan += kPeakWindow;
            scan_done = window_end != 0 || scan >= count;
            __syncthreads();  // the queue is written; window_end and offsets may be reused
        }
        if (q_head >= q_len) break;  // nothing left to resolve
        // one block of queued survivors, in sorted order
        const int e = q_head + tid;
        const bool valid = e < q_len;
        const int id = valid ? q_id[e] : -1;
        float c[kPeakMaxDim];
#pragma unroll
        for (int j = 0; j < kPeakMaxDim; ++j)
            if (j < dim) c[j] = valid ? q_c[e * dim + j] : 0.0f;
        const int n = n_taken;
        const bool ok = valid && !clashes(c, n);
        alive[tid] = ok;
#pragma unroll
        for (int j = 0; j < kPeakMaxDim; ++j)
            if (j < dim) block_c[tid * dim + j] = c[j];
        __syncthreads();
        // the later candidates of the block that clash with this one
        unsigned long long row[kPeakBlock / 64] = {0ull, 0ull, 0ull, 0ull};
        if (ok)
            for (int q = tid + 1; q < kPeakBlock; ++q)
                if (alive[q] && peak_d2(c, block_c + q * dim, dim) < sep2) row[q >> 6] |= 1ull << (q & 63);
        for (int w = 0; w < kPeakBlock / 64; ++w) conflict[tid][w] = row[w];
        __syncthreads();
        // wave 0 walks the block in order on scalar bit masks: the first live candidate is
        // taken and every later one that clashes with it leaves, so the walk takes one turn
        // per taken peak, however many near-duplicates the block holds
        if (wave == 0) {
            unsigned long long rows[kPeakBlock / 64][kPeakBlock / 64];
            unsigned long long live[kPeakBlock / 64];
#pragma unroll
            for (int slot = 0; slot < kPeakBlock / 64; ++slot) {
#pragma unroll
                for (int w = 0; w < kPeakBlock / 64; ++w) rows[slot][w] = conflict[slot * 64 + lane][w];
                live[slot] = __ballot(alive[slot * 64 + lane] != 0);
            }
            unsigned long long mine[kPeakBlock / 64] = {0ull, 0ull, 0ull, 0ull};
            int k = n;
#pragma unroll
            for (int slot = 0; slot < kPeakBlock / 64; ++slot) {
                while (live[slot] != 0ull && k < peaks) {
                    const int l = __builtin_ctzll(live[slot]);
                    live[slot] &= live[slot] - 1ull;
                    mine[slot] |= 1ull << l;
                    ++k;
#pragma unroll
                    for (int w = slot; w < kPeakBlock / 64; ++w) live[w] &= ~readlane64(rows[slot][w], l);
                }
            }
            if (lane == 0) {
#pragma unroll
                for (int w = 0; w < kPeakBlock / 64; ++w) chosen[w] = mine[w];
                n_taken = k;
                finished = k >= peaks;
            }
        }
        __syncthreads();
        // every taken candidate writes its own slot: its rank among this block's choices
        const unsigned long long bits = chosen[tid >> 6];
        if ((bits >> lane) & 1ull) {
            int k = n + __popcll(bits & ((1ull << lane) - 1ull));
            for (int w = 0; w < (tid >> 6); ++w) k += __popcll(chosen[w]);
            out[k] = id;
            float* dst = k < lds_cap ? taken_lds + k * dim : spill + (size_t)(k - lds_cap) * dim;
#pragma unroll
            for (int j = 0; j < kPeakMaxDim; ++j)
                if (j < dim) dst[j] = c[j];
            if (sep2 > 0.0f && grid) {
                const int next = atomicExch(&head[peak_bucket(cell(c, 0), cell(c, 1), cell(c, 2))], k);
                if (k < kPeakChainLds)
                    chain_lds[k] = next;
                else
                    chain[k] = next;
            }
        }
        __syncthreads();
        if (finished) break;
        q_head = min(q_head + kPeakBlock, q_len);
    }
    for (int k = n_taken + tid; k < peaks; k += kPeakBlock) out[k] = -1;
}

size_t peaks_workspace(int count, int peaks, int dim) {
    size_t sort_bytes = 0;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(
        nullptr, sort_bytes, (const float*)nullptr, (float*)nullptr, (const int*)nullptr, (int*)nullptr, count);
    if (e != hipSuccess) return 0;
    const size_t arrays = 4 * (size_t)count * sizeof(float);
    return 512 + arrays + sizeof(float) * (size_t)peaks * (dim + 1) + sort_bytes;
}

// -H^{-1} g (or -g if not a descent direction), masked, tangent to unit normals and
// scaled into the trust radius (search.slang:66-91).
__global__ __launch_bounds__(128) void direction_kernel(const float* __restrict__ inv, const float* __restrict__ grad,
                                                        const float* __restrict__ nrm, const float* __restrict__ mask,
                                                        int count, int width, float radius, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float d[6];
    float descent = 0.0f;
    for (int j = 0; j < width; ++j) {
        float v = 0.0f;
        for (int k = 0; k < width; ++k) v -= inv[((size_t)i * width + j) * width + k] * grad[(size_t)i * width + k];
        d[j] = v;
        descent += v * grad[(size_t)i * width + j];
    }
    for (int j = 0; j < width; ++j) d[j] = (descent >= 0.0f ? -grad[(size_t)i * width + j] : d[j]) * mask[j];
    if (width == 6) {
        float proj = 0.0f;
        for (int j = 0; j < 3; ++j) proj += d[j + 3] * nrm[3 * i + j];
        for (int j = 0; j < 3; ++j) d[j + 3] -= proj * nrm[3 * i + j];
    }
    float len2 = 0.0f;
    for (int j = 0; j < width; ++j) len2 += d[j] * d[j];
    const float scale = fminf(1.0f, radius / fmaxf(sqrtf(len2), 1e-30f));
    for (int j = 0; j < width; ++j) out[(size_t)i * width + j] = d[j] * scale;
}

// Step a fraction along the direction: centers clamped to the box where trainable,
// normals renormalized when they move (search.slang:93-109). controls [4, 6]:
// per-coordinate scale, trainable mask, lower and upper bounds.
__global__ __launch_bounds__(128) void trial_kernel(const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                    const float* __restrict__ dir, const float* __restrict__ ctl,
                                                    int count, int width, float fraction, float* __restrict__ oc,
                                                    float* __restrict__ on) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float n[3];
    float len2 = 0.0f;
    for (int j = 0; j < 3; ++j) {
        const float c = ctr[3 * i + j] + fraction * ctl[j] * dir[(size_t)i * width + j];
        oc[3 * i + j] = ctl[6 + j] > 0.0f ? fminf(fmaxf(c, ctl[12 + j]), ctl[18 + j]) : ctr[3 * i + j];
        n[j] = nrm[3 * i + j] + (width == 6 ? fraction * ctl[j + 3] * dir[(size_t)i * width + j + 3] : 0.0f);
        len2 += n[j] * n[j];
    }
    const float inv = width == 6 ? 1.0f / sqrtf(len2) : 1.0f;
    for (int j = 0; j < 3; ++j) on[3 * i + j] = width == 6 ? n[j] * inv : nrm[3 * i + j];
}

hipError_t launch_range_gate(const float* pts, const float* pos, const int* intervals, int count, int apertures,
                             int bins, float spacing, int required, int* out, hipStream_t st) {
    hipLaunchKernelGGL(range_gate_kernel, dim3((count + 255) / 256), dim3(256), 0, st, pts, pos, intervals, count,
                       apertures, bins, spacing, required, out);
    return hipGetLastError();
}

hipError_t launch_peaks(const float* scores, const float* ctr, int count, int peaks, int dim, float separation,
                        int* out, void* workspace, size_t bytes, hipStream_t st) {
    char* w = static_cast<char*>(workspace);
    float* keys_in = reinterpret_cast<float*>(w);
    float* keys = keys_in + count;
    int* ids_in = reinterpret_cast<int*>(keys + count);
    int* order = ids_in + count;
    float* spill = reinterpret_cast<float*>(order + count);
    int* chain = reinterpret_cast<int*>(spill + (size_t)peaks * dim);
    char* temp = reinterpret_cast<char*>(chain + peaks);
    temp += (256 - reinterpret_cast<uintptr_t>(temp) % 256) % 256;
    size_t temp_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp_bytes, keys_in, keys, ids_in, order,
                                                                count, 0, 32, st);
    if (e != hipSuccess) return e;
    if (temp + temp_bytes > w + bytes) return hipErrorInvalidValue;
    if (count > 0) {
        hipLaunchKernelGGL(peak_keys, dim3((count + 255) / 256), dim3(256), 0, st, scores, count, keys_in, ids_in);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys, ids_in, order, count, 0,
                                                         32, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(peaks_kernel, dim3(1), dim3(kPeakBlock), 0, st, keys, order, ctr, count, peaks, dim,
                       separation * separation, spill, chain, out);
    return hipGetLastError();
}

hipError_t launch_direction(const float* inv, const float* grad, const float* nrm, const float* mask, int count,
                            int width, float radius, float* out, hipStream_t st) {
    hipLaunchKernelGGL(direction_kernel, dim3((count + 127) / 128), dim3(128), 0, st, inv, grad, nrm, mask, count,
                       width, radius, out);
    return hipGetLastError();
}

hipError_t launch_trial(const float* ctr, const float* nrm, const float* dir, const float* ctl, int count, int width,
                        float fraction, float* oc, float* on, hipStream_t st) {
    hipLaunchKernelGGL(trial_kernel, dim3((count + 127) / 128), dim3(128), 0, st, ctr, nrm, dir, ctl, count, width,
                       fraction, oc, on);
    return hipGetLastError();
}

}  // namespace dsplat

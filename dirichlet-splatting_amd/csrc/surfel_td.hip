// Surfel splat forward and VJP in the sample domain, on MFMA (reference semantics:
// dsplat/kernels/surfel.slang:87-311, bindings dsplat/kernels/surfel.py:58-137).
//
// The Dirichlet kernel is the finite sum D(d) = (1/S) sum_{n<S} e^{-2 pi i n d / N}, so with
// v = e^{2 pi i k0 / N} per pair and the pair coefficient c = w amp rot:
//   out[a, k]  = (1/S) sum_n e^{-2 pi i n k / N} X[a, n],     X[a, n] = sum_s c_{a,s} v_{a,s}^n
//   g0[a, s]   = sum_k conj(D(k - k0)) u[a, k]  = (1/S) sum_n w^n U[a, n]
//   g1[a, s]   = sum_k conj(D'(k - k0)) u[a, k] = (2 pi i / (N S)) sum_n n w^n U[a, n]
// with w = conj(v) and U[a, n] = sum_{k<K} e^{2 pi i n k / N} u[a, k]. Splitting n = 16 j + i,
// v^n = Z_j W_i (W_i = v^i, Z_j = v^{16 j}), the sums over surfels (forward) and over i (VJP)
// are small complex GEMMs per aperture whose operands are generated per pair: they run on
// v_mfma_f32_16x16x4f32 / v_mfma_f32_32x32x2f32 at full FP32 precision. The sums are exact
// (no lobe special case, no reciprocal per bin); the only approximations are the float32
// roundings of the generated powers (binary-tree products from an exactly rounded e^{2 pi i m/N}
// table times a small rotation, error growing like log2(n) ulps).
//
// Forward (td_render_kernel<J>, S = 16 J samples): a workgroup owns 16 apertures. Surfels
// stream in chunks of 16; every thread generates one pair (Z_0..Z_{J-1}, c W_0..c W_15)
// into LDS, then each wave accumulates X of its 4 apertures as [2J x 16] real tiles:
//   rows (j, part), cols i, K = (surfel, part): Re X += Zr (cW)r - Zi (cW)i, Im X += Zi (cW)r + Zr (cW)i.
// The DFT to bins (a [16 x S] x [S x K] complex GEMM with twiddles from the LDS table) runs
// in the same workgroup, so X never leaves the chip. This FP32 form serves 256 samples and the
// JVP; 128 samples use td_render_h16_kernel (FP16 hi + lo products, symmetric split, below).
//
// VJP (td_vjp_kernel<J>): a workgroup owns 128 surfels (32 per wave, lane % 32 = surfel) and a
// range of apertures; U (the DFT of the cotangent, from the dft_rows_kernel pre-pass) is
// staged per aperture in the MFMA A-operand layout. The two lane halves generate two pairs
// (apertures a and a + 1, same surfel) and swap half of their W operands, so every pair is
// generated once. With the symmetric split n = 16 j + 7.5 +- (i + 0.5), i < 8 (vjp_operands_kernel),
// per aperture M0[s, j] = sum_i W_i U+ + conj(W_i) U- and M1 the same with n U[n] form one
// 32-row tile per 8 j on v_mfma_f32_32x32x16_f16 (FP16 hi + lo operands, K = 16: Re and Im of
// the 8 W_i = w^(i + 1/2)); P0 = sum_j Y_j M0 with Y_j = w^(16 j + 7.5), P1 likewise.
// The 8 chained gradients accumulate in registers over the aperture range (no cross-lane
// reduction per pair) and are written as one partial per (surfel, range), summed in float64
// by vjp_reduce.
#include "dsplat_device.h"
#include "dsplat_internal.h"
#include "dsplat_td.h"

namespace dsplat {

// dft16 for 8 rows: the 16 MFMA rows take the two halves of the samples, row i the samples
// p < Lin / 2 of X row i % 8 (i < 8) or p + Lin / 2 (i >= 8), so the K loop is half as long.
// The second half's result is turned by e^{sign 2 pi i (Lin / 2) q / N} (Re and Im of a bin
// sit in neighbouring lanes: one quad swap) and added to the first half's across the lane
// halves (rows i and i + 8 are lanes l and l + 32). Lin must be a multiple of 4.
__device__ __forceinline__ void dft8_split(const float2* X, int ldx, int Lin, const float2* E, int N, float sign,
                                           float scale, int Lout, float2* out, int ldo, int rows, int wave,
                                           int lane) {
    const int kk = lane >> 4, x = lane & 15;
    const int pin = kk & 1, dp = kk >> 1, pout = x & 1;
    const int half = Lin >> 1;
    const float* Xf = reinterpret_cast<const float*>(X);
    const float* xrow = Xf + 2 * ((x & 7) * ldx + dp + (x >> 3) * half) + pin;
    for (int ct = wave; ct * 8 < Lout; ct += 4) {
        const int q = ct * 8 + (x >> 1);
        const f32x4 d = dft_tile(xrow, E, (dp * q) & (N - 1), (2 * q) & (N - 1), N, scale, sign, pin, pout, half / 2);
        const float2 rho = E[(half * q) & (N - 1)];
        const float rr = rho.x, ri = sign * rho.y;
        float res[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // a lane shuffle rather than update_dpp: hipcc merged four update_dpp calls on the
            // elements of one MFMA result into one, giving every element element 0's neighbour
            const float other = __shfl_xor(d[r], 1, 64);
            const float re = pout ? other : d[r], im = pout ? d[r] : other;
            float turned = pout ? fmaf(re, ri, im * rr) : fmaf(re, rr, -im * ri);
            float recv = turned;
            swap_halves(turned, recv);  // lanes 0..31: recv = turned of lane + 32
            res[r] = d[r] + recv;
        }
        if (kk < 2) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 4 * kk + r;
                if (row < rows) reinterpret_cast<float*>(out + (size_t)row * ldo + q)[pout] = res[r];
            }
        }
    }
}




// ---------------------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------------------
template <int J, bool JVP>
struct TdRec {
    // floats per pair record. Forward: Z re[J], Z im[J], (c W) re[16], (c W) im[16]. JVP (two
    // coefficient sets sharing the powers): (e1 Z) re[J], (e1 Z) im[J], (e2 Z) re[J],
    // (e2 Z) im[J], W re[16], W im[16]. Plus 2 floats so the record stride is 2 mod 4 dwords:
    // the 8-byte writes of 16 consecutive producer lanes then fall on distinct LDS banks.
    static constexpr int sets = JVP ? 2 : 1;
    static constexpr int used = 2 * J * sets + 32;
    static constexpr int size = used + 2;
    static constexpr int shift = 0;  // dwords added after every group of 4 records
    __device__ static int at(int t) { return t * size + shift * (t >> 2); }
    static size_t bytes(int n) { return sizeof(float) * ((size_t)n * size + shift * (size_t)(n / 4)); }
};

// Per-surfel inputs of a producer, loaded one chunk ahead.
template <bool JVP>
struct SurfIn {
    float cx, cy, cz, nx, ny, nz;
    float2 w;
    float dc[JVP ? 3 : 1], dn[JVP ? 3 : 1];
    float2 dw;
};

template <bool JVP>
__device__ __forceinline__ SurfIn<JVP> load_surfel(int s, const float* __restrict__ ctr,
                                                   const float* __restrict__ nrm, const float2* __restrict__ w,
                                                   const float* __restrict__ dctr, const float* __restrict__ dnrm,
                                                   const float2* __restrict__ dw) {
    SurfIn<JVP> in;
    in.cx = ctr[3 * s];
    in.cy = ctr[3 * s + 1];
    in.cz = ctr[3 * s + 2];
    in.nx = nrm[3 * s];
    in.ny = nrm[3 * s + 1];
    in.nz = nrm[3 * s + 2];
    in.w = w[s];
    if constexpr (JVP) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            in.dc[d] = dctr[3 * s + d];
            in.dn[d] = dnrm[3 * s + d];
        }
        in.dw = dw[s];
    }
    return in;
}

// Producer: the record of pair (aperture p, surfel s) into `mine` (LDS). Forward
// (JVP = false): e = c = w amp rot. JVP (surfel.slang:146-221): dy = sum_s e1 D + e2 D', so
// X = X1 - (2 pi i n / N) X2 with X1 = sum_s e1 v^n, X2 = sum_s e2 v^n,
// e1 = rot ((damp + i amp dphase) w + amp dw), e2 = -amp k0pm dr rot w; both share Z.
template <int J, bool JVP, bool FRESNEL>
__device__ __forceinline__ void td_produce(const Chirp& c, const float2* E, bool ok, float px, float py, float pz,
                                           const SurfIn<JVP>& in, float* mine) {
    constexpr int SETS = TdRec<J, JVP>::sets;
    float2 coef[SETS];
    const Pair q = pair_geometry_t<FRESNEL>(c, px, py, pz, in.cx, in.cy, in.cz, in.nx, in.ny, in.nz);
    const float amp = ok ? q.amp : 0.0f;
    if constexpr (JVP) {
        const float dr = -(q.ux * in.dc[0] + q.uy * in.dc[1] + q.uz * in.dc[2]);
        const float dcos = ((q.cosine * q.ux - in.nx) * in.dc[0] + (q.cosine * q.uy - in.ny) * in.dc[1] +
                            (q.cosine * q.uz - in.nz) * in.dc[2]) * q.inv_r +
                           q.ux * in.dn[0] + q.uy * in.dn[1] + q.uz * in.dn[2];
        const float damp = -amp * dr * q.inv_r + (ok ? q.amp_cos : 0.0f) * dcos;
        coef[0] = cmul(q.rot, cmul(make_float2(damp, amp * q.rate * dr), in.w) + amp * in.dw);
        coef[SETS - 1] = (-amp * c.k0pm_f * dr) * cmul(q.rot, in.w);
    } else {
        coef[0] = amp * cmul(in.w, q.rot);
    }
    const float2 v = rotation(c, E, q.lobe, q.delta, 1);
    const float2 z1 = rotation(c, E, q.lobe, q.delta, 16);
    const float2 v2 = csq(v);
    if constexpr (JVP) {
        // e Z_j = e z1^j for both sets by chains from e; W_i = v^i once, two chains by v^2
#pragma unroll
        for (int e = 0; e < SETS; ++e) {
            float2 a0 = coef[e];
            float* dst = mine + 2 * J * e;
#pragma unroll
            for (int j = 0; j < J; j += 2) {
                const float2 a1 = cmul(a0, z1);
                *reinterpret_cast<float2*>(dst + j) = make_float2(a0.x, a1.x);
                *reinterpret_cast<float2*>(dst + J + j) = make_float2(a0.y, a1.y);
                a0 = cmul(a1, z1);
            }
        }
        float2 w0 = make_float2(1.f, 0.f), w1 = v;
        float* dst = mine + 2 * J * SETS;
#pragma unroll
        for (int i = 0; i < 16; i += 2) {
            *reinterpret_cast<float2*>(dst + i) = make_float2(w0.x, w1.x);
            *reinterpret_cast<float2*>(dst + 16 + i) = make_float2(w0.y, w1.y);
            w0 = i == 0 ? v2 : cmul(w0, v2);
            w1 = cmul(w1, v2);
        }
        return;
    }
    // Z_j = z1^j by a binary product tree
    float2 Z[J];
    Z[0] = make_float2(1.f, 0.f);
    Z[1] = z1;
#pragma unroll
    for (int j = 2; j < J; ++j) Z[j] = (j & 1) ? cmul(Z[j - 1], z1) : csq(Z[j >> 1]);
#pragma unroll
    for (int j = 0; j < J; j += 2) {
        *reinterpret_cast<float2*>(mine + j) = make_float2(Z[j].x, Z[j + 1].x);
        *reinterpret_cast<float2*>(mine + J + j) = make_float2(Z[j].y, Z[j + 1].y);
    }
    // c W_i = c v^i, two interleaved chains by v^2
    float2 W[16];
    W[0] = coef[0];
    W[1] = cmul(coef[0], v);
#pragma unroll
    for (int i = 2; i < 16; ++i) W[i] = cmul(W[i - 2], v2);
    float* dst = mine + 2 * J;
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
        *reinterpret_cast<float2*>(dst + i) = make_float2(W[i].x, W[i + 1].x);
        *reinterpret_cast<float2*>(dst + 16 + i) = make_float2(W[i].y, W[i + 1].y);
    }
}

// Consumer: wave cw accumulates its 4 apertures over one chunk of SC surfels on
// v_mfma_f32_16x16x4f32 (A[i][k] -> lane i + 16 k, B[k][j] -> lane j + 16 k,
// C[4 (lane / 16) + r][lane % 16] in register r). The four real products of the complex one
// are kept apart, so no operand needs a sign: K index kk = surfel 4 t + kk; row x =
// (j = x / 2, Z part x % 2) of the set's A block; column tile (set e, W part p), column x = i.
// Then Re X = C(re)[(j, re)] - C(im)[(j, im)],  Im X = C(im)[(j, re)] + C(re)[(j, im)].
// Forward: A = Z, B = c W; JVP: A = e Z per set, B = W shared.
template <int J, int SETS, int T, int SC>
__device__ __forceinline__ void td_consume(const float* rec, int cw, int lane, f32x4 (&acc)[SETS][2][4][T]) {
    constexpr int REC = TdRec<J, SETS == 2>::size;
    const int kk = lane >> 4, x = lane & 15;
    const int a_off = (x & 1) * J + (x >> 1);  // Z part x % 2 of j = x / 2
    const int b_off = 2 * J * SETS + x;        // W re of i = x; im at + 16
    constexpr int GROUP = 4 * REC + TdRec<J, SETS == 2>::shift;  // dwords per group of 4 records
    const float* base = rec + kk * REC + (4 * cw) * (SC / 4) * GROUP;
    // all operands of a K step are read before its products, so the LDS reads overlap
#pragma unroll
    for (int t = 0; t < SC / 4; ++t) {
        float bv[4][2], av[4][T][SETS];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float* r = base + (b * (SC / 4) + t) * GROUP;
            bv[b][0] = r[b_off];
            bv[b][1] = r[b_off + 16];
#pragma unroll
            for (int tt = 0; tt < T; ++tt)
#pragma unroll
                for (int e = 0; e < SETS; ++e) av[b][tt][e] = r[a_off + 2 * J * e + 8 * tt];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int tt = 0; tt < T; ++tt)
#pragma unroll
                for (int e = 0; e < SETS; ++e)
#pragma unroll
                    for (int p = 0; p < 2; ++p)
                        acc[e][p][b][tt] =
                            __builtin_amdgcn_mfma_f32_16x16x4f32(av[b][tt][e], bv[b][p], acc[e][p][b][tt], 0, 0, 0);
    }
}

// X tiles of consumer wave cw -> LDS rows [4 cw + b][S + 1] float2. Register r of lane (g, i)
// is row 4 g + r = (j = 2 g + r / 2, Z part r % 2) of sample 16 j + i.
template <int J, int SETS, int T>
__device__ __forceinline__ void td_store_x(const Chirp& c, float2* X, int cw, int lane,
                                           const f32x4 (&acc)[SETS][2][4][T]) {
    const int ldx = c.S + 1;
    const int g = lane >> 4, i = lane & 15;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int tt = 0; tt < T; ++tt)
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const int j = 8 * tt + 2 * g + (r >> 1);
                if (j >= J) continue;
                const int n = 16 * j + i;
                float2 xs[SETS];
#pragma unroll
                for (int e = 0; e < SETS; ++e)
                    xs[e] = make_float2(acc[e][0][b][tt][r] - acc[e][1][b][tt][r + 1],
                                        acc[e][1][b][tt][r] + acc[e][0][b][tt][r + 1]);
                float2 val = xs[0];
                if constexpr (SETS == 2) {
                    // - (2 pi i n / N) X2
                    const float k = c.two_pi_N * (float)n;
                    val.x += k * xs[1].y;
                    val.y -= k * xs[1].x;
                }
                X[(4 * cw + b) * ldx + n] = val;
            }
}

// Single-buffer form (JVP): all 4 waves produce one pair each per chunk, then consume.
template <int J, bool JVP, bool FRESNEL>
__global__ __launch_bounds__(256) void td_render_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                        const float* __restrict__ ctr,
                                                        const float* __restrict__ nrm,
                                                        const float2* __restrict__ w,
                                                        const float* __restrict__ dctr,
                                                        const float* __restrict__ dnrm,
                                                        const float2* __restrict__ dw, int S,
                                                        float2* __restrict__ out) {
    constexpr int AT = kTdApertures, SC = kTdChunk, SETS = TdRec<J, JVP>::sets;
    constexpr int T = (2 * J + 15) / 16;
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);  // [N]
    float* rec = smem + 2 * c.N;                  // AT * SC records; later X [AT][S + 1] float2

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int a0 = blockIdx.x * AT;
    fill_twiddles(E, c.N, tid, 256);
    const int pal = tid >> 4, psp = tid & 15;
    const int pa = min(a0 + pal, A - 1);
    const bool p_ok = a0 + pal < A;
    const float px = pos[3 * pa], py = pos[3 * pa + 1], pz = pos[3 * pa + 2];

    f32x4 acc[SETS][2][4][T];
#pragma unroll
    for (int e = 0; e < SETS; ++e)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int tt = 0; tt < T; ++tt) acc[e][p][b][tt] = f32x4{0.f, 0.f, 0.f, 0.f};

    SurfIn<JVP> in = load_surfel<JVP>(min(psp, S - 1), ctr, nrm, w, dctr, dnrm, dw);
    __syncthreads();  // twiddle table written
    // Wave w produces the records of apertures 4 w .. 4 w + 3 and is their only consumer, so
    // the chunk loop needs no workgroup barrier: the LDS executes one wave's operations in
    // issue order, and the wave barriers keep the compiler from moving the record writes of a
    // chunk across the reads of the previous one (or the reads ahead of the writes).
    for (int s0 = 0; s0 < S; s0 += SC) {
        const int s = s0 + psp;
        const SurfIn<JVP> next = load_surfel<JVP>(min(s + SC, S - 1), ctr, nrm, w, dctr, dnrm, dw);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        td_produce<J, JVP, FRESNEL>(c, E, p_ok && s < S, px, py, pz, in, rec + TdRec<J, JVP>::at(tid));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        td_consume<J, SETS, T, SC>(rec, wave, lane, acc);
        in = next;
    }
    __syncthreads();
    float2* X = reinterpret_cast<float2*>(rec);
    td_store_x<J, SETS, T>(c, X, wave, lane, acc);
    __syncthreads();
    dft16(X, c.S + 1, c.S, E, c.N, -1.0f, c.inv_S, c.K, out + (size_t)a0 * c.K, c.K, min(AT, A - a0), wave, lane);
}

// ---------------------------------------------------------------------------------------
// Forward on FP16 hi + lo MFMA (128 samples)
// ---------------------------------------------------------------------------------------
// FP32 products issue on the SIMD cycles that the pair generation needs for its VALU work;
// FP16 products co-issue with it. td_render_h16_kernel (forward, and the JVP with two
// coefficient sets) forms the sums over surfels on
// v_mfma_f32_16x16x32_f16 as hi hi + hi lo + lo hi of FP16 splits with FP32 accumulation.
//  * Samples are split symmetrically, n = 16 j + 7 + i with i in [-7, 8]: v^n = Y_j W_i with
//    Y_j = c v^(16 j + 7) and W_i = v^i, and W_-i = conj(W_i). With the four real sums
//    P1 = sum Re Y Re W, P2 = sum Im Y Im W, P3 = sum Im Y Re W, P4 = sum Re Y Im W over the
//    surfels, X[16 j + 7 + i] = (P1 - P2) + i (P3 + P4) and X[16 j + 7 - i] = (P1 + P2) +
//    i (P3 - P4): nine W columns (W_0 = 1) give all 16 samples of a row j, and a pair
//    supplies 32 generated values (Y_0..7, W_1..8) for 128 complex products.
//  * A workgroup is 4 waves x 2 apertures; surfels stream in chunks of 32. Lane (h, m) of
//    wave w generates pair (aperture 2 w + m / 16, chunk surfel t + 16 h), t = m % 16.
//    Its values go in row pairs (rowbase, rowbase + 1) through one v_permlane32_swap, after
//    which half h holds row rowbase + h of surfels t and t + 16, and one packed split gives
//    the hi and lo words [row][t] = (surfel t, surfel t + 16).
//  * Rows per aperture: 0..15 Y (row 2 j + part), 16 + x Re W, 32 + x Im W for the B column
//    x: W_(x+1) for x < 8, W_0 (constant) at x = 8, zero for x > 8 (the constant columns are
//    held once per workgroup, H16Layout). Consumer lane (x, g) reads the words 4 g .. 4 g + 3
//    of its row: K slot
//    8 g + 2 u + b of A and B alike is surfel 4 g + u + 16 b.
//  * Row R holds 16 hi words then 16 lo words, in the other order when (R / 2) is odd, and
//    word group g sits at group g ^ ((R / 4) % 4): the 16 lanes of a 128-bit read cover all
//    64 banks. The second aperture's rows start 16 words further, so the 64 lanes of one
//    producer dword write fall on distinct banks (a paired hi / lo write still meets the other
//    aperture's words: the 2-way conflicts that remain).
//  * Y is scaled by 2^sh with one exponent per wave, so that the largest value of the chunks
//    so far lies below 2^13.5 (FP16 keeps 22 bits of every value); sh only decreases, the
//    accumulators are rescaled by 2^(sh_new - sh_old) when it does, and X = acc 2^-sh.
//  * Records are wave-private, so the chunk loop has no workgroup barrier.
constexpr int kH16Apertures = 8;
constexpr int kH16Chunk = 32;
// Rows are numbered virtually: SETS x 16 Y rows, then the Re W and Im W tiles of 16 rows each
// (tile column x = row - first row of the tile). Columns 0..7 of both tiles change with the
// pair; columns 8..15 are constant (W_0 = 1 at Re column 8, zeros elsewhere) and sit once per
// workgroup in a shared block of 16 slots. Per aperture, slots hold the Y rows, Re W columns
// 0..7 and Im W columns 0..7. A slot has the parity of its virtual row, so the bank pattern of
// h16_word (which depends on the row mod 16 and the slot parity) is the virtual row's.
template <bool JVP>
struct H16Layout {
    static constexpr int sets = JVP ? 2 : 1;
    static constexpr int rw = 16 * sets;               // first virtual Re W row
    static constexpr int slots = rw + 16;              // per aperture
    static constexpr int ap_words = slots * 32 + 16;   // [slot][hi | lo 16 words] + bank shift
    static constexpr int wave_words = 2 * ap_words;    // per wave: two apertures
    static constexpr int const_words = 16 * 32;        // per workgroup
    static constexpr int pairs = slots / 2;            // row pairs written per chunk
    // first virtual row of producer row pair p: Y rows 2 p, Re W rw + 2 (p - 8 sets), Im W
    // rw + 16 + 2 (p - 8 sets - 4)
    __host__ __device__ static constexpr int rowbase(int p) { return p < 8 * sets + 4 ? 2 * p : 2 * p + 8; }
    // per-aperture slot of a changing virtual row
    __host__ __device__ static constexpr int slot(int R) { return R < rw + 16 ? R : R - 8; }
};

// word offset of the (hi | lo) half of row R, word group g, word u within the group
__host__ __device__ constexpr int h16_word(int R, int lo, int g, int u) {
    return R * 32 + 16 * (((R >> 1) & 1) ^ lo) + 4 * (g ^ ((R >> 2) & 3)) + u;
}

// Maximum of x over the wave (uniform result): quad, half-row and row mirrors by DPP, then
// the four rows by readlane.
__device__ __forceinline__ int wave_max_i(int x) {
    x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));   // quad_perm [1, 0, 3, 2]
    x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));   // quad_perm [2, 3, 0, 1]
    x = max(x, __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false));  // row_half_mirror
    x = max(x, __builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false));  // row_mirror
    return max(max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
               max(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

// biased exponent of max(|Re z|, |Im z|)
__device__ __forceinline__ int cexp_bits(float2 z) {
    return (int)(max(__float_as_uint(z.x) & 0x7fffffffu, __float_as_uint(z.y) & 0x7fffffffu) >> 23);
}

// Forward (JVP = false) and JVP (JVP = true: two coefficient sets e1, e2 with one W set,
// X = X1 - (2 pi i n / N) X2 as in td_render_kernel) on the layout above.
template <bool FRESNEL, bool JVP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void td_render_h16_kernel(
    const Chirp c, const float* __restrict__ pos, int A, const float* __restrict__ ctr,
    const float* __restrict__ nrm, const float2* __restrict__ w, const float* __restrict__ dctr,
    const float* __restrict__ dnrm, const float2* __restrict__ dw, int S, float2* __restrict__ out) {
    using L = H16Layout<JVP>;
    constexpr int SETS = L::sets, RW = 16 * SETS;  // RW: first Re W row
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);  // [N]
    float* region = smem + 2 * c.N;               // 4 waves of records; later X [8][S + 1] float2
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, m = lane & 31;
    uint32_t* rec = reinterpret_cast<uint32_t*>(region) + wave * L::wave_words;
    uint32_t* cst = reinterpret_cast<uint32_t*>(region) + 4 * L::wave_words;  // shared constant slots
    const int a0 = blockIdx.x * kH16Apertures;
    const int pb = m >> 4, t = m & 15;
    const int a = a0 + 2 * wave + pb;
    const bool p_ok = a < A;
    const int pa = min(a, A - 1);
    const float px = pos[3 * pa], py = pos[3 * pa + 1], pz = pos[3 * pa + 2];
    fill_twiddles(E, c.N, tid, 256);
    // constant slots: Re W columns 8..15 (slots 0..7), Im W columns 8..15 (slots 8..15); only the
    // hi words of Re W_0 (column 8) are nonzero: (1.0, 1.0) in FP16
    for (int i = tid; i < L::const_words; i += 256) {
        const int sl = i >> 5, wd = i & 31;
        const int row = RW + 8 + (sl < 8 ? sl : 8 + sl);  // its virtual row
        const bool hi_half = (wd < 16) != (((row >> 1) & 1) != 0);  // hi words come second when (row / 2) is odd
        cst[i] = (sl == 0 && hi_half) ? 0x3C003C00u : 0u;
    }

    // producer: lane word offsets for the four swizzle keys (row pair p uses key (rowbase / 4) % 4)
    int wofs[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) wofs[k] = pb * L::ap_words + h * 32 + 4 * ((t >> 2) ^ k) + (t & 3);
    // consumer: row 16 e + x (A of set e), RW + x (B re) or RW + 16 + x (B im) of aperture b, word
    // group g; B columns x >= 8 come from the shared constant slots
    const int x = lane & 15, g = lane >> 4;
    int bofs[2][2];  // [tile][hi, lo]: word offset from the aperture base (x < 8) or from cst (x >= 8)
#pragma unroll
    for (int tl = 0; tl < 2; ++tl)
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) {
            const int R = RW + 16 * tl + x;
            const int sl = x < 8 ? RW + 8 * tl + x : 8 * tl + x - 8;
            bofs[tl][hl] = h16_word(R, hl, g, 0) + (sl - R) * 32;
        }

    f32x4 acc[SETS][2][2];
#pragma unroll
    for (int e = 0; e < SETS; ++e)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int p = 0; p < 2; ++p) acc[e][b][p] = f32x4{0.f, 0.f, 0.f, 0.f};
    int e_run[SETS], sh[SETS];
#pragma unroll
    for (int e = 0; e < SETS; ++e) {
        e_run[e] = 0;
        sh[e] = 127;
    }

    SurfIn<JVP> in = load_surfel<JVP>(min(t + 16 * h, S - 1), ctr, nrm, w, dctr, dnrm, dw);
    __syncthreads();  // twiddle table and constant rows written
    for (int s0 = 0; s0 < S; s0 += kH16Chunk) {
        const int s = s0 + t + 16 * h;
        const SurfIn<JVP> next = load_surfel<JVP>(min(s + kH16Chunk, S - 1), ctr, nrm, w, dctr, dnrm, dw);
        const Pair q = pair_geometry_t<FRESNEL>(c, px, py, pz, in.cx, in.cy, in.cz, in.nx, in.ny, in.nz);
        const bool ok = p_ok && s < S;
        const float amp = ok ? q.amp : 0.0f;
        float2 coef[SETS];
        if constexpr (JVP) {
            // surfel.slang:146-221: e1 = rot ((damp + i amp dphase) w + amp dw), e2 = -amp k0pm dr rot w
            const float dr = -(q.ux * in.dc[0] + q.uy * in.dc[1] + q.uz * in.dc[2]);
            const float dcos = ((q.cosine * q.ux - in.nx) * in.dc[0] + (q.cosine * q.uy - in.ny) * in.dc[1] +
                                (q.cosine * q.uz - in.nz) * in.dc[2]) * q.inv_r +
                               q.ux * in.dn[0] + q.uy * in.dn[1] + q.uz * in.dn[2];
            const float damp = -amp * dr * q.inv_r + (ok ? q.amp_cos : 0.0f) * dcos;
            coef[0] = cmul(q.rot, cmul(make_float2(damp, amp * q.rate * dr), in.w) + amp * in.dw);
            coef[SETS - 1] = (-amp * c.k0pm_f * dr) * cmul(q.rot, in.w);
        } else {
            coef[0] = amp * cmul(in.w, q.rot);
        }
        // per set a wave-uniform running exponent of max |Re e|, |Im e| (biased); |Y_j| = |e|
#pragma unroll
        for (int e = 0; e < SETS; ++e) {
            const int e_chunk = wave_max_i(cexp_bits(coef[e]));
            if (e_chunk > e_run[e]) {
                e_run[e] = e_chunk;
                const int sh_new = min(139 - e_chunk, 127);
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int p = 0; p < 2; ++p)
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc[e][b][p][r] = ldexpf(acc[e][b][p][r], sh_new - sh[e]);
                sh[e] = sh_new;
            }
            coef[e] = make_float2(ldexpf(coef[e].x, sh[e]), ldexpf(coef[e].y, sh[e]));
        }
        const float2 v = rotation(c, E, q.lobe, q.delta, 1);
        const float2 z1 = rotation(c, E, q.lobe, q.delta, 16);
        float val[2 * L::pairs];
        {
            // W_1..8 by a product tree, Y_j = e W_7 z1^j by two chains over z1^2
            float2 W[9];
            W[1] = v;
#pragma unroll
            for (int i = 2; i <= 8; ++i) W[i] = (i & 1) ? cmul(W[i - 1], v) : csq(W[i >> 1]);
            const float2 z2 = csq(z1);
#pragma unroll
            for (int e = 0; e < SETS; ++e) {
                float2 Y[8];
                Y[0] = cmul(coef[e], W[7]);
                Y[1] = cmul(Y[0], z1);
#pragma unroll
                for (int j = 2; j < 8; ++j) Y[j] = cmul(Y[j - 2], z2);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    val[16 * e + 2 * j] = Y[j].x;
                    val[16 * e + 2 * j + 1] = Y[j].y;
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                val[RW + i] = W[i + 1].x;
                val[RW + 8 + i] = W[i + 1].y;
            }
        }
        // the previous chunk's reads were issued before these writes (in-order LDS per wave)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int p = 0; p < L::pairs; ++p) {
            float x0 = val[2 * p], x1 = val[2 * p + 1];
            swap_halves(x0, x1);  // half h: row rowbase(p) + h of surfels t (x0) and t + 16 (x1)
            f16x2 hi, lo;
            split_f16_pair<false>(x0, x1, hi, lo);
            const int R = L::rowbase(p);  // even: h only moves the row by one
            uint32_t* dst = rec + wofs[(R >> 2) & 3] + (L::slot(R) - R) * 32;
            dst[h16_word(R, 0, 0, 0) - 4 * ((R >> 2) & 3)] = __builtin_bit_cast(uint32_t, hi);
            dst[h16_word(R, 1, 0, 0) - 4 * ((R >> 2) & 3)] = __builtin_bit_cast(uint32_t, lo);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        f16x8 op[2][2][SETS + 2];  // [aperture][hi, lo][A of each set, B re, B im]
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int hl = 0; hl < 2; ++hl) {
#pragma unroll
                for (int o = 0; o < SETS; ++o)
                    op[b][hl][o] = __builtin_bit_cast(
                        f16x8, *reinterpret_cast<const uint4*>(rec + b * L::ap_words + h16_word(16 * o + x, hl, g, 0)));
#pragma unroll
                for (int tl = 0; tl < 2; ++tl)
                    op[b][hl][SETS + tl] = __builtin_bit_cast(
                        f16x8, *reinterpret_cast<const uint4*>((x < 8 ? rec + b * L::ap_words : cst) + bofs[tl][hl]));
            }
        // the accumulators interleaved, so no product waits for the one before it
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int e = 0; e < SETS; ++e)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int p = 0; p < 2; ++p)
                        acc[e][b][p] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            op[b][k == 2][e], op[b][k == 1][SETS + p], acc[e][b][p], 0, 0, 0);
        in = next;
    }
    __syncthreads();  // every wave's records consumed: the region becomes X
    float2* X = reinterpret_cast<float2*>(region);
    const int ldx = c.S + 1;
    {
        // register r of lane (g, x) is row 4 g + r = (j = 2 g + r / 2, Y part r % 2) of column
        // x, i.e. W_(x + 1) for x < 8 and W_0 for x = 8
        const int i = x < 8 ? x + 1 : 0;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const int n0 = 16 * (2 * g + (r >> 1)) + 7;
                float2 plus[SETS], minus[SETS];
#pragma unroll
                for (int e = 0; e < SETS; ++e) {
                    const f32x4& ac0 = acc[e][b][0];
                    const f32x4& ac1 = acc[e][b][1];
                    const float P1 = ac0[r], P3 = ac0[r + 1], P4 = ac1[r], P2 = ac1[r + 1];
                    plus[e] = make_float2(ldexpf(P1 - P2, -sh[e]), ldexpf(P3 + P4, -sh[e]));
                    minus[e] = make_float2(ldexpf(P1 + P2, -sh[e]), ldexpf(P3 - P4, -sh[e]));
                }
                float2 vp = plus[0], vm = minus[0];
                if constexpr (JVP) {
                    // - (2 pi i n / N) X2
                    const float kp = c.two_pi_N * (float)(n0 + i), km = c.two_pi_N * (float)(n0 - i);
                    vp = make_float2(vp.x + kp * plus[1].y, vp.y - kp * plus[1].x);
                    vm = make_float2(vm.x + km * minus[1].y, vm.y - km * minus[1].x);
                }
                float2* row = X + (2 * wave + b) * ldx;
                if (x <= 8) row[n0 + i] = vp;
                if (x < 7) row[n0 - i] = vm;
            }
    }
    __syncthreads();
    dft8_split(X, ldx, c.S, E, c.N, -1.0f, c.inv_S, c.K, out + (size_t)a0 * c.K, c.K, min(kH16Apertures, A - a0),
               wave, lane);
}

// ---------------------------------------------------------------------------------------
// DFT of rows: out[a][q] = scale sum_{p<Lin} e^{sign 2 pi i p q / N} in[a][p], 16 rows per
// workgroup (the VJP pre-pass U = DFT of the cotangent, sign +1).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dft_rows_kernel(const float2* __restrict__ in, int rows, int Lin, int N,
                                                       float sign, float scale, int Lout,
                                                       float2* __restrict__ out) {
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);
    float2* X = E + N;
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * 16;
    const int ldx = Lin + 1;
    fill_twiddles(E, N, tid, 256);
    for (int i = tid; i < 16 * Lin; i += 256) {
        const int r = i / Lin, p = i - r * Lin;
        X[r * ldx + p] = r0 + r < rows ? in[(size_t)(r0 + r) * Lin + p] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    dft16(X, ldx, Lin, E, N, sign, scale, Lout, out + (size_t)r0 * Lout, Lout, min(16, rows - r0), tid >> 6,
          tid & 63);
}



// A operands of the VJP GEMM for v_mfma_f32_32x32x16_f16 (A[i][8 h + e] -> lane i + 32 h,
// element e), as FP16 hi + lo splits of the per-aperture scaled values, built once per
// aperture from U: [A][T tiles][2 (hi, lo)][64 lanes] x 8 halves, and the scale exponent e_a
// per aperture (values were multiplied by 2^-e_a, max |value| below 2^12).
// The samples are split symmetrically with half-integer powers: n = 16 j + 7.5 +- (i + 0.5),
// i < 8, so sum_n w^n U[n] = sum_j Y_j M_j with Y_j = w^(16 j + 7.5), W_i = w^(i + 0.5) and
// M_j = sum_i W_i U+ + conj(W_i) U- = sum_i Re W_i (U+ + U-) + i Im W_i (U+ - U-), where
// U+ = U[16 j + 8 + i], U- = U[16 j + 7 - i] (times n for mat 1). Row -> (mat, j, part) as
// the accumulator layout (td_row); K = 2 i + k, k = 0 (B = Re W_i) or 1 (B = Im W_i):
// (re) Re S, (im) Im S for k = 0; (re) -Im D, (im) Re D for k = 1 (S = U+ + U-, D = U+ - U-).
template <int J>
__global__ __launch_bounds__(64 * (J / 8)) void vjp_operands_kernel(const float2* __restrict__ U, int S,
                                                                    uint4* __restrict__ out,
                                                                    float* __restrict__ scale_exp) {
    constexpr int T = J / 8;
    __shared__ float red[T];
    const int a = blockIdx.x, tid = threadIdx.x, tile = tid >> 6, L = tid & 63;
    const int row = L & 31, h = L >> 5;
    int mat, j;
    td_row<J>(tile, row >> 3, (row >> 2) & 1, row & 3, &mat, &j);
    const int part = row & 1;
    float v[8];
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int K = 8 * h + e, i = K >> 1, k = K & 1;
        const int np = 16 * j + 8 + i, nm = 16 * j + 7 - i;
        float2 up = U[(size_t)a * S + np], um = U[(size_t)a * S + nm];
        if (mat) {
            up = (float)np * up;
            um = (float)nm * um;
        }
        const float2 sum = up + um, dif = up - um;
        const float x = k == 0 ? (part ? sum.y : sum.x) : (part ? dif.x : -dif.y);
        v[e] = x;
        mx = fmaxf(mx, fabsf(x));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (L == 0) red[tile] = mx;
    __syncthreads();
    for (int t = 0; t < T; ++t) mx = fmaxf(mx, red[t]);
    int ex = 0;
    if (mx > 0.0f) frexpf(mx, &ex);  // mx < 2^ex
    ex -= 12;
    if (tid == 0) scale_exp[a] = (float)ex;
    const float sc = ldexpf(1.0f, -ex);
    f16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float x = v[e] * sc;
        const _Float16 xh = (_Float16)x;
        hi[e] = xh;
        lo[e] = (_Float16)(x - (float)xh);
    }
    uint4* dst = out + (((size_t)a * T + tile) * 2) * 64 + L;
    dst[0] = __builtin_bit_cast(uint4, hi);
    dst[64] = __builtin_bit_cast(uint4, lo);
}

// The VJP pre-pass for 128 samples in one launch: 16 cotangent rows are taken to the sample
// domain in LDS (dft16, sign +1), then each wave builds one aperture's operand tiles from
// them as vjp_operands_kernel<8> does, so U never goes through HBM.
__global__ __launch_bounds__(256) void vjp_prep_kernel(const float2* __restrict__ up, int A, int K, int N, int S,
                                                       uint4* __restrict__ out, float* __restrict__ scale_exp) {
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);  // [N]
    float2* Xin = E + N;                          // [16][K + 1]
    float2* U = Xin + 16 * (K + 1);               // [16][S + 1]
    const int tid = threadIdx.x, wave = tid >> 6, L = tid & 63;
    const int r0 = blockIdx.x * 16;
    fill_twiddles(E, N, tid, 256);
    for (int i = tid; i < 16 * K; i += 256) {
        const int r = i / K, p = i - r * K;
        Xin[r * (K + 1) + p] = r0 + r < A ? up[(size_t)(r0 + r) * K + p] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    dft16(Xin, K + 1, K, E, N, 1.0f, 1.0f, S, U, S + 1, 16, wave, L);
    __syncthreads();
    const int row = L & 31, h = L >> 5;
    int mat, j;
    td_row<8>(0, row >> 3, (row >> 2) & 1, row & 3, &mat, &j);
    const int part = row & 1;
    for (int q = wave; q < 16; q += 4) {
        const int a = r0 + q;  // uniform over the wave
        if (a >= A) break;
        const float2* Ua = U + q * (S + 1);
        float v[8];
        float mx = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int Kx = 8 * h + e, i = Kx >> 1, k = Kx & 1;
            const int np = 16 * j + 8 + i, nm = 16 * j + 7 - i;
            float2 upl = Ua[np], umi = Ua[nm];
            if (mat) {
                upl = (float)np * upl;
                umi = (float)nm * umi;
            }
            const float2 sum = upl + umi, dif = upl - umi;
            const float x = k == 0 ? (part ? sum.y : sum.x) : (part ? dif.x : -dif.y);
            v[e] = x;
            mx = fmaxf(mx, fabsf(x));
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
        int ex = 0;
        if (mx > 0.0f) frexpf(mx, &ex);  // mx < 2^ex
        ex -= 12;
        if (L == 0) scale_exp[a] = (float)ex;
        const float sc = ldexpf(1.0f, -ex);
        f16x8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float x = v[e] * sc;
            const _Float16 xh = (_Float16)x;
            hi[e] = xh;
            lo[e] = (_Float16)(x - (float)xh);
        }
        uint4* dst = out + ((size_t)a * 2) * 64 + L;
        dst[0] = __builtin_bit_cast(uint4, hi);
        dst[64] = __builtin_bit_cast(uint4, lo);
    }
}

// One workgroup: 4 waves x 32 surfels, an aperture range staged G apertures at a time. Each
// step covers apertures (a, a + 1): the two lane halves generate one pair each, then
// 3 x 2 x T FP16 hi + lo MFMAs per aperture (hi hi + hi lo + lo hi, FP32 accumulation), the
// contraction over j and the chain rule.
template <int J, bool FRESNEL>
__global__ __launch_bounds__(256) void td_vjp_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                     const float* __restrict__ ctr,
                                                     const float* __restrict__ nrm,
                                                     const float2* __restrict__ w, int S,
                                                     const uint4* __restrict__ Aexp,
                                                     const float* __restrict__ scale_exp, int per_range,
                                                     float* __restrict__ partial) {
    constexpr int T = J / 8;  // 32-row tiles per aperture
    constexpr int G = kTdVjpGroup / T;
    constexpr int PER = T * 2 * 64;  // uint4 per aperture
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);         // [N]
    float* Sc = reinterpret_cast<float*>(E + c.N);       // [G] scales 2^e_a
    uint4* Aop = reinterpret_cast<uint4*>(Sc + 16);      // [G][T][2][64]

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, m = lane & 31;
    // the surfel tiles of one aperture range run on one XCD and read its operands through one L2
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, range = lin / gridDim.x;
    const int s = tile * kTdVjpSurfels + 32 * wave + m;
    const bool s_ok = s < S;
    const int sc = min(s, S - 1);
    const int a_begin = range * per_range, a_end = min(A, a_begin + per_range);
    fill_twiddles(E, c.N, tid, 256);

    const float cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
    const float nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
    const float2 ws = w[sc];
    float gacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

    for (int g0 = a_begin; g0 < a_end; g0 += G) {
        __syncthreads();
        // Stage the group's operands. Apertures past the range take a clamped aperture's
        // (finite) operands: their pairs have amp = amp_cos = 0, which zeroes every gradient term.
        stage_to_lds<G * PER / 256>(Aop, tid, [&](int e) {
            return Aexp + (size_t)min(g0 + e / PER, A - 1) * PER + e % PER;
        });
        if (tid < G) Sc[tid] = ldexpf(1.0f, (int)scale_exp[min(g0 + tid, A - 1)]);
        __syncthreads();
        for (int ga = 0; ga < G; ga += 2) {
            VjpStage<J> cur;
            {
                const int a = g0 + ga + h;  // this lane's pair: (a, s)
                vjp_stage<J, FRESNEL>(c, E, pos, min(a, A - 1), a < a_end && s_ok, cx, cy, cz, nx, ny, nz, cur);
            }
            f32x16 acc[2][T];
#pragma unroll
            for (int ap = 0; ap < 2; ++ap)
#pragma unroll
                for (int tile = 0; tile < T; ++tile)
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[ap][tile][r] = 0.f;
#pragma unroll
            for (int ap = 0; ap < 2; ++ap)
#pragma unroll
                for (int tile = 0; tile < T; ++tile) {
                    const uint4* src = Aop + ((ga + ap) * T + tile) * 2 * 64 + lane;
                    const f16x8 ah = __builtin_bit_cast(f16x8, src[0]);
                    const f16x8 al = __builtin_bit_cast(f16x8, src[64]);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bh[ap], acc[ap][tile], 0, 0, 0);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bl[ap], acc[ap][tile], 0, 0, 0);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, cur.Bh[ap], acc[ap][tile], 0, 0, 0);
                }
            // Contraction P_mat = sum_j Z_j M_mat[j]. Half h holds the rows of j in its set J_h
            // for both apertures; one half swap of (Z_{j of half 0}, Z_{j of half 1}) gives each
            // half Z_{j in J_h} of aperture a and of aperture a + 1.
            float2 P[2][2];  // [aperture a / a + 1][mat]
#pragma unroll
            for (int x = 0; x < 2; ++x) P[x][0] = P[x][1] = make_float2(0.f, 0.f);
#pragma unroll
            for (int tile = 0; tile < T; ++tile)
#pragma unroll
                for (int rr = 0; rr < 16; rr += 2) {
                    int mat, j0, j1;
                    td_row<J>(tile, rr >> 2, 0, rr & 3, &mat, &j0);
                    td_row<J>(tile, rr >> 2, 1, rr & 3, &mat, &j1);
                    float2 za0 = cur.Z[j0], za1 = cur.Z[j1];
                    swap_halves(za0.x, za1.x);
                    swap_halves(za0.y, za1.y);
                    P[0][mat] = cmac(P[0][mat], za0, make_float2(acc[0][tile][rr], acc[0][tile][rr + 1]));
                    P[1][mat] = cmac(P[1][mat], za1, make_float2(acc[1][tile][rr], acc[1][tile][rr + 1]));
                }
            // undo the per-aperture operand scales, then each half completes the sums of its own
            // pair's aperture (a + h) with the other half's share: after the swap, P[0] + P[1]
            // is that total on both halves.
            const float s0 = Sc[ga], s1 = Sc[ga + 1];
#pragma unroll
            for (int mat = 0; mat < 2; ++mat) {
                P[0][mat] = s0 * P[0][mat];
                P[1][mat] = s1 * P[1][mat];
                swap_halves(P[0][mat].x, P[1][mat].x);
                swap_halves(P[0][mat].y, P[1][mat].y);
            }
            const float2 P0 = P[0][0] + P[1][0];
            const float2 P1 = P[0][1] + P[1][1];
            // g0 = P0 / S, g1 = (2 pi i / (N S)) P1; chained as surfel.slang:294-302.
            const float2 g0 = c.inv_S * P0;
            const float2 g1 = (c.two_pi_N * c.inv_S) * make_float2(-P1.y, P1.x);
            const float2 z0 = cmulc(cur.rot, g0), zz1 = cmulc(cur.rot, g1);
            const float2 cw = cmulc(ws, z0);
            const float along = cmulc(ws, zz1).x * c.k0pm_f;
            const float gr = -cw.x * cur.amp * cur.inv_r + cur.amp * (cw.y * cur.rate - along);
            const float gcos = cw.x * cur.amp_cos;
            gacc[0] += -gr * cur.ux + gcos * (cur.cosine * cur.ux - nx) * cur.inv_r;
            gacc[1] += -gr * cur.uy + gcos * (cur.cosine * cur.uy - ny) * cur.inv_r;
            gacc[2] += -gr * cur.uz + gcos * (cur.cosine * cur.uz - nz) * cur.inv_r;
            gacc[3] += gcos * cur.ux;
            gacc[4] += gcos * cur.uy;
            gacc[5] += gcos * cur.uz;
            gacc[6] += cur.amp * z0.x;
            gacc[7] += cur.amp * z0.y;
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float x = gacc[j], y = gacc[j];
        swap_halves(x, y);
        gacc[j] = x + y;
    }
    if (h == 0 && s_ok) {
        float4* dst = reinterpret_cast<float4*>(partial + ((size_t)s * gridDim.y + range) * 8);
        dst[0] = make_float4(gacc[0], gacc[1], gacc[2], gacc[3]);
        dst[1] = make_float4(gacc[4], gacc[5], gacc[6], gacc[7]);
    }
}

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
bool td_supported(const Chirp& c) {
    // S = 16 J samples with J in {8, 16}; the Z_1 rotation needs 16 <= N / 4.
    return (c.S == 128 || c.S == 256) && c.N >= 64 && (c.N & (c.N - 1)) == 0 && c.K % 8 == 0;
}

bool use_td(const Chirp& c) { return c.form == 0 && td_supported(c); }

template <int J, bool JVP>
static hipError_t launch_td_render_j(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                     const float2* w, const float* dctr, const float* dnrm, const float2* dw, int S,
                                     float2* out, hipStream_t st) {
    const size_t rec = TdRec<J, JVP>::bytes(kTdApertures * kTdChunk);
    const size_t xbuf = sizeof(float2) * (size_t)kTdApertures * (c.S + 1);
    const size_t smem = sizeof(float2) * c.N + (rec > xbuf ? rec : xbuf);
    const bool fres = c.fresnel_n > 1.0f;
    hipError_t e = allow_lds(fres ? (const void*)td_render_kernel<J, JVP, true> : (const void*)td_render_kernel<J, JVP, false>,
                             smem);
    if (e != hipSuccess) return e;
    const dim3 grid((A + kTdApertures - 1) / kTdApertures);
    if (fres)
        hipLaunchKernelGGL((td_render_kernel<J, JVP, true>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, dctr,
                           dnrm, dw, S, out);
    else
        hipLaunchKernelGGL((td_render_kernel<J, JVP, false>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, dctr,
                           dnrm, dw, S, out);
    return hipGetLastError();
}

template <bool JVP>
static hipError_t launch_td_render_h16(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                       const float2* w, const float* dctr, const float* dnrm, const float2* dw, int S,
                                       float2* out, hipStream_t st) {
    const size_t region = sizeof(uint32_t) * (4 * (size_t)H16Layout<JVP>::wave_words + H16Layout<JVP>::const_words);
    const size_t xbuf = sizeof(float2) * kH16Apertures * (size_t)(c.S + 1);
    const size_t smem = sizeof(float2) * c.N + (region > xbuf ? region : xbuf);
    const bool fres = c.fresnel_n > 1.0f;
    const void* fn = fres ? (const void*)td_render_h16_kernel<true, JVP> : (const void*)td_render_h16_kernel<false, JVP>;
    hipError_t e = allow_lds(fn, smem);
    if (e != hipSuccess) return e;
    const dim3 grid((A + kH16Apertures - 1) / kH16Apertures);
    if (fres)
        hipLaunchKernelGGL((td_render_h16_kernel<true, JVP>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, dctr,
                           dnrm, dw, S, out);
    else
        hipLaunchKernelGGL((td_render_h16_kernel<false, JVP>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, dctr,
                           dnrm, dw, S, out);
    return hipGetLastError();
}

hipError_t launch_td_render(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                            const float2* w, int S, float2* out, hipStream_t st) {
    if (c.S == 128) return launch_td_render_h16<false>(c, pos, A, ctr, nrm, w, nullptr, nullptr, nullptr, S, out, st);
    if (c.S == 256) return launch_td_render_j<16, false>(c, pos, A, ctr, nrm, w, nullptr, nullptr, nullptr, S, out, st);
    return hipErrorInvalidValue;
}

hipError_t launch_td_jvp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, const float* dctr, const float* dnrm, const float2* dw, int S, float2* out,
                         hipStream_t st) {
    if (c.S == 128) return launch_td_render_h16<true>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
    if (c.S == 256) return launch_td_render_j<16, true>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
    return hipErrorInvalidValue;
}

TdVjpShape td_vjp_shape(int A, int S) {
    TdVjpShape v;
    v.tiles = (S + kTdVjpSurfels - 1) / kTdVjpSurfels;
    // enough workgroups to fill 256 CUs several times; ranges are multiples of the staging group
    const int max_ranges = (A + kTdVjpGroup - 1) / kTdVjpGroup;
    int ranges = (2048 + v.tiles - 1) / v.tiles;
    ranges = ranges < 1 ? 1 : (ranges > max_ranges ? max_ranges : ranges);
    int per = (A + ranges - 1) / ranges;
    v.per_range = (per + kTdVjpGroup - 1) / kTdVjpGroup * kTdVjpGroup;
    v.ranges = (A + v.per_range - 1) / v.per_range;
    return v;
}

// Workspace: partials [S][ranges][8] floats | U [A][S] complex (256 samples only: at 128
// samples vjp_prep_kernel keeps U in LDS) | A operands [A][T][2][64] uint4 | scale exponents
// [A] floats, each part 256-byte aligned.
struct TdVjpWorkspace {
    size_t partial, u, ops, scales, total;
};

static TdVjpWorkspace td_vjp_layout(const Chirp& c, int A, int S) {
    const TdVjpShape v = td_vjp_shape(A, S);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    TdVjpWorkspace w;
    w.partial = 0;
    w.u = up(sizeof(float) * 8 * (size_t)S * v.ranges);
    w.ops = w.u + (c.S == 128 ? 0 : up(sizeof(float2) * (size_t)A * c.S));
    w.scales = w.ops + up(sizeof(uint4) * (size_t)A * (c.S / 128) * 2 * 64);
    w.total = w.scales + up(sizeof(float) * (size_t)A);
    return w;
}

size_t td_vjp_workspace(const Chirp& c, int A, int S) {
    return vjp_moments_supported(c) ? vjp_moments_workspace(A, S) : td_vjp_layout(c, A, S).total;
}

template <int J>
static hipError_t launch_td_vjp_j(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                  const float2* w, int S, const float2* up, void* ws, float* gc, float* gn,
                                  float* gw, hipStream_t st) {
    constexpr int T = J / 8;
    const TdVjpShape v = td_vjp_shape(A, S);
    const TdVjpWorkspace lay = td_vjp_layout(c, A, S);
    char* base = static_cast<char*>(ws);
    float* partial = reinterpret_cast<float*>(base + lay.partial);
    float2* U = reinterpret_cast<float2*>(base + lay.u);
    uint4* ops = reinterpret_cast<uint4*>(base + lay.ops);
    float* scales = reinterpret_cast<float*>(base + lay.scales);
    if (J == 8) {
        // U = DFT of the cotangent rows (K bins -> S samples, sign +1) and the operand tiles in one launch
        const size_t smem = sizeof(float2) * (c.N + 16 * (size_t)(c.K + 1) + 16 * (size_t)(c.S + 1));
        hipError_t e = allow_lds((const void*)vjp_prep_kernel, smem);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(vjp_prep_kernel, dim3((A + 15) / 16), dim3(256), smem, st, up, A, c.K, c.N, c.S, ops,
                           scales);
    } else {
        // U = DFT of the cotangent rows: K bins -> S samples, sign +1
        const size_t smem = sizeof(float2) * (c.N + 16 * (size_t)(c.K + 1));
        hipError_t e = allow_lds((const void*)dft_rows_kernel, smem);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(dft_rows_kernel, dim3((A + 15) / 16), dim3(256), smem, st, up, A, c.K, c.N, 1.0f, 1.0f,
                           c.S, U);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(vjp_operands_kernel<J>, dim3(A), dim3(64 * T), 0, st, U, c.S, ops, scales);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t smem = sizeof(float2) * c.N + sizeof(float) * 16 +
                        sizeof(uint4) * (size_t)(kTdVjpGroup / T) * T * 2 * 64;
    const void* fn = c.fresnel_n > 1.0f ? (const void*)td_vjp_kernel<J, true> : (const void*)td_vjp_kernel<J, false>;
    e = allow_lds(fn, smem);
    if (e != hipSuccess) return e;
    if (c.fresnel_n > 1.0f)
        hipLaunchKernelGGL((td_vjp_kernel<J, true>), dim3(v.tiles, v.ranges), dim3(256), smem, st, c, pos, A, ctr,
                           nrm, w, S, ops, scales, v.per_range, partial);
    else
        hipLaunchKernelGGL((td_vjp_kernel<J, false>), dim3(v.tiles, v.ranges), dim3(256), smem, st, c, pos, A, ctr,
                           nrm, w, S, ops, scales, v.per_range, partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_vjp_reduce(partial, S, v.ranges, gc, gn, gw, st);
}

hipError_t launch_dft_rows_up(const Chirp& c, const float2* in, int rows, float2* out, hipStream_t st) {
    const size_t smem = sizeof(float2) * (c.N + 16 * (size_t)(c.K + 1));
    hipError_t e = allow_lds((const void*)dft_rows_kernel, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dft_rows_kernel, dim3((rows + 15) / 16), dim3(256), smem, st, in, rows, c.K, c.N, 1.0f, 1.0f,
                       c.S, out);
    return hipGetLastError();
}

hipError_t launch_td_vjp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, int S, const float2* up, void* ws, float* gc, float* gn, float* gw,
                         hipStream_t st) {
    if (vjp_moments_supported(c)) return launch_vjp_moments(c, pos, A, ctr, nrm, w, S, up, ws, gc, gn, gw, st);
    if (c.S == 128) return launch_td_vjp_j<8>(c, pos, A, ctr, nrm, w, S, up, ws, gc, gn, gw, st);
    if (c.S == 256) return launch_td_vjp_j<16>(c, pos, A, ctr, nrm, w, S, up, ws, gc, gn, gw, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------
// Certificate scan in the sample domain (S = 128): per (candidate, aperture)
//   corr = amp conj(rot) g0,  g0 = (1/S) sum_n w^n R[a, n]   (R = the residual row taken to
//   the sample domain), and norm = amp^2 f(k0), f(k0) = sum_{k<K} |D(k - k0)|^2.
// f is a trigonometric polynomial of degree S - 1 in v = e^{2 pi i k0 / N}:
//   f = 2 Re sum_{m=0}^{S-1} phi_m v^m,  phi_0 = K / (2 S),  phi_m = (S - m) E_m / S^2,
//   E_m = sum_{k<K} e^{-2 pi i m k / N} = 1 - i cot(pi m / N) for odd m, 0 for even m (K = N / 2),
// so with psi = conj(phi) it is 2 Re sum_n w^n psi_n: the same product as g0 (split as the VJP's)
// with psi in place of R. The 32 accumulator rows per aperture hold both: rows of mat 0 are
// the residual's (scaled by 2^-e_a per aperture), rows of mat 1 the fixed psi (scaled by 2^10).
// ---------------------------------------------------------------------------------------

struct TdScanSet {
    int count;
    const float* positions[kMaxScans];
    int apertures[kMaxScans];
    int ap_off[kMaxScans + 1];       // prefix sums of apertures
    int range_first[kMaxScans + 1];  // prefix sums of aperture ranges
    int per_range;
};

// Operands as vjp_operands_kernel<J> (symmetric half-integer split, one K half, T = J / 8
// tiles) for the rows of both mats: S = U+ + U-, D = U+ - U- of the residual (mat 0, scaled
// per aperture) and of psi (mat 1, the fixed norm-polynomial rows).
template <int J>
__global__ __launch_bounds__(64 * (J / 8)) void scan_operands_kernel(const float2* __restrict__ U, int S, int N,
                                                                     int K, uint4* __restrict__ out,
                                                                     float* __restrict__ scale_exp) {
    constexpr int T = J / 8;
    __shared__ float red[T];
    const int a = blockIdx.x, tid = threadIdx.x, tile = tid >> 6, L = tid & 63;
    const int row = L & 31, h = L >> 5;
    int mat, j;
    td_row<J>(tile, row >> 3, (row >> 2) & 1, row & 3, &mat, &j);
    const int part = row & 1;
    // psi_m = conj(phi_m), scaled by 2^kScanPsiExp
    auto psi = [&](int m) {
        double re = 0.0, im = 0.0;
        if (m == 0) {
            re = (double)K / (2.0 * S);
        } else if (m & 1) {
            double sn, cs;
            sincospi((double)m / (double)N, &sn, &cs);
            const double w = (double)(S - m) / ((double)S * S);
            re = w;
            im = w * cs / sn;
        }
        return make_float2((float)ldexp(re, kScanPsiExp), (float)ldexp(im, kScanPsiExp));
    };
    float v[8];
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int Kx = 8 * h + e, i = Kx >> 1, k = Kx & 1;
        const int np = 16 * j + 8 + i, nm = 16 * j + 7 - i;
        const float2 up = mat == 0 ? U[(size_t)a * S + np] : psi(np);
        const float2 um = mat == 0 ? U[(size_t)a * S + nm] : psi(nm);
        const float2 sum = up + um, dif = up - um;
        const float x = k == 0 ? (part ? sum.y : sum.x) : (part ? dif.x : -dif.y);
        v[e] = x;
        if (mat == 0) mx = fmaxf(mx, fabsf(x));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (L == 0) red[tile] = mx;
    __syncthreads();
    for (int t = 0; t < T; ++t) mx = fmaxf(mx, red[t]);  // the residual rows of every tile
    int ex = 0;
    if (mx > 0.0f) frexpf(mx, &ex);
    ex -= 12;
    if (tid == 0) scale_exp[a] = (float)ex;
    const float sc = mat == 0 ? ldexpf(1.0f, -ex) : 1.0f;
    f16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float x = v[e] * sc;
        const _Float16 xh = (_Float16)x;
        hi[e] = xh;
        lo[e] = (_Float16)(x - (float)xh);
    }
    uint4* dst = out + (((size_t)a * T + tile) * 2) * 64 + L;
    dst[0] = __builtin_bit_cast(uint4, hi);
    dst[64] = __builtin_bit_cast(uint4, lo);
}

// grid (candidate tiles of 128, aperture ranges of every scan back to back); partial
// [C][ranges][3] = (corr.re, corr.im, norm) of the range.
template <int J, bool FRESNEL>
__global__ __launch_bounds__(256) void scan_td_kernel(const Chirp c, const TdScanSet set,
                                                      const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                      int C, const uint4* __restrict__ ops,
                                                      const float* __restrict__ scale_exp,
                                                      float* __restrict__ partial) {
    constexpr int T = J / 8;  // 32-row tiles per aperture
    constexpr int G = kTdVjpGroup / T, PER = T * 2 * 64;
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);     // [N]
    float* Sc = reinterpret_cast<float*>(E + c.N);   // [G] scales 2^e_a
    uint4* Aop = reinterpret_cast<uint4*>(Sc + 16);  // [G][T][2][64]

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, m = lane & 31;
    const int s = blockIdx.x * kScanCandidates + 32 * wave + m;
    const bool s_ok = s < C;
    const int sc = min(s, C - 1);
    const int range = blockIdx.y, nrange = gridDim.y;
    int k = 0;
    while (k + 1 < set.count && range >= set.range_first[k + 1]) ++k;
    const int A = set.apertures[k];
    const float* __restrict__ pos = set.positions[k];
    const uint4* __restrict__ kops = ops + (size_t)set.ap_off[k] * PER;
    const float* __restrict__ kexp = scale_exp + set.ap_off[k];
    const int a_begin = (range - set.range_first[k]) * set.per_range;
    const int a_end = min(A, a_begin + set.per_range);
    fill_twiddles(E, c.N, tid, 256);

    const float cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
    const float nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
    const float psi_scale = ldexpf(1.0f, -kScanPsiExp);
    float acc_re = 0.f, acc_im = 0.f, acc_n = 0.f;

    for (int g0 = a_begin; g0 < a_end; g0 += G) {
        __syncthreads();
        // past the range: a clamped aperture's operands, zeroed by amp = 0 of those pairs
        stage_to_lds<G * PER / 256>(Aop, tid, [&](int e) {
            return kops + (size_t)min(g0 + e / PER, A - 1) * PER + e % PER;
        });
        if (tid < G) Sc[tid] = ldexpf(1.0f, (int)kexp[min(g0 + tid, A - 1)]);
        __syncthreads();
        for (int ga = 0; ga < G; ga += 2) {
            VjpStage<J> cur;
            {
                const int a = g0 + ga + h;  // this lane's pair: (a, s)
                vjp_stage<J, FRESNEL>(c, E, pos, min(a, A - 1), a < a_end && s_ok, cx, cy, cz, nx, ny, nz, cur);
            }
            f32x16 acc[2][T];
#pragma unroll
            for (int ap = 0; ap < 2; ++ap)
#pragma unroll
                for (int tile = 0; tile < T; ++tile) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) acc[ap][tile][r] = 0.f;
                    const uint4* src = Aop + ((ga + ap) * T + tile) * 2 * 64 + lane;
                    const f16x8 ah = __builtin_bit_cast(f16x8, src[0]);
                    const f16x8 al = __builtin_bit_cast(f16x8, src[64]);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bh[ap], acc[ap][tile], 0, 0, 0);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bl[ap], acc[ap][tile], 0, 0, 0);
                    acc[ap][tile] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, cur.Bh[ap], acc[ap][tile], 0, 0, 0);
                }
            float2 P[2][2];  // [aperture a / a + 1][mat 0: residual, mat 1: psi]
#pragma unroll
            for (int x = 0; x < 2; ++x) P[x][0] = P[x][1] = make_float2(0.f, 0.f);
#pragma unroll
            for (int tile = 0; tile < T; ++tile)
#pragma unroll
                for (int rr = 0; rr < 16; rr += 2) {
                    int mat, j0, j1;
                    td_row<J>(tile, rr >> 2, 0, rr & 3, &mat, &j0);
                    td_row<J>(tile, rr >> 2, 1, rr & 3, &mat, &j1);
                    float2 za0 = cur.Z[j0], za1 = cur.Z[j1];
                    swap_halves(za0.x, za1.x);
                    swap_halves(za0.y, za1.y);
                    P[0][mat] = cmac(P[0][mat], za0, make_float2(acc[0][tile][rr], acc[0][tile][rr + 1]));
                    P[1][mat] = cmac(P[1][mat], za1, make_float2(acc[1][tile][rr], acc[1][tile][rr + 1]));
                }
            P[0][0] = Sc[ga] * P[0][0];
            P[1][0] = Sc[ga + 1] * P[1][0];
#pragma unroll
            for (int mat = 0; mat < 2; ++mat) {
                swap_halves(P[0][mat].x, P[1][mat].x);
                swap_halves(P[0][mat].y, P[1][mat].y);
            }
            const float2 g0 = c.inv_S * (P[0][0] + P[1][0]);
            const float f = 2.0f * psi_scale * (P[0][1].x + P[1][1].x);
            const float2 corr = cmulc(cur.rot, g0);
            acc_re = fmaf(cur.amp, corr.x, acc_re);
            acc_im = fmaf(cur.amp, corr.y, acc_im);
            acc_n = fmaf(cur.amp * cur.amp, f, acc_n);
        }
    }
    {
        float x = acc_re, y = acc_re;
        swap_halves(x, y);
        acc_re = x + y;
        x = acc_im, y = acc_im;
        swap_halves(x, y);
        acc_im = x + y;
        x = acc_n, y = acc_n;
        swap_halves(x, y);
        acc_n = x + y;
    }
    if (h == 0 && s_ok) {
        float* dst = partial + ((size_t)s * nrange + range) * 3;
        dst[0] = acc_re;
        dst[1] = acc_im;
        dst[2] = acc_n;
    }
}

// out[scan][s][3] = sum of partial[s][r][3] over the scan's ranges, float64, fixed order.
__global__ __launch_bounds__(256) void scan_td_reduce(const float* __restrict__ partial, int C, int nrange,
                                                      const TdScanSet set, float* __restrict__ out) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int k = blockIdx.y;
    if (row >= C) return;
    const int r0 = set.range_first[k], r1 = set.range_first[k + 1];
    for (int f = 0; f < 3; ++f) {
        double acc = 0.0;
        for (int r = r0 + lane; r < r1; r += 64) acc += partial[((size_t)row * nrange + r) * 3 + f];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) out[((size_t)k * C + row) * 3 + f] = (float)acc;
    }
}

bool td_scan_supported(const Chirp& c) { return use_td(c); }

static TdScanSet td_scan_set(const ScanSet& set, int chunk) {
    TdScanSet t;
    t.count = set.count;
    t.per_range = chunk < kScanMinRange ? kScanMinRange : (chunk + kTdVjpGroup - 1) / kTdVjpGroup * kTdVjpGroup;
    t.ap_off[0] = 0;
    t.range_first[0] = 0;
    for (int i = 0; i < set.count; ++i) {
        t.positions[i] = set.positions[i];
        t.apertures[i] = set.apertures[i];
        t.ap_off[i + 1] = t.ap_off[i] + set.apertures[i];
        t.range_first[i + 1] = t.range_first[i] + (set.apertures[i] + t.per_range - 1) / t.per_range;
    }
    return t;
}

// Workspace: partials [C][ranges][3] (ranges at the smallest range length) | U [sum A][S]
// complex | operands [sum A][128] uint4 | scale exponents [sum A], each 256-byte aligned.
struct TdScanWorkspace {
    size_t partial, u, ops, scales, total;
};

static TdScanWorkspace td_scan_layout(const Chirp& c, const ScanSet& set, int C) {
    const TdScanSet t = td_scan_set(set, kScanMinRange);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t apertures = t.ap_off[set.count];
    TdScanWorkspace w;
    w.partial = 0;
    w.u = up(sizeof(float) * 3 * (size_t)C * t.range_first[set.count]);
    w.ops = w.u + up(sizeof(float2) * apertures * c.S);
    w.scales = w.ops + up(sizeof(uint4) * apertures * 128 * (c.S / 128));  // T = S / 128 tiles
    w.total = w.scales + up(sizeof(float) * apertures);
    return w;
}

size_t td_scan_workspace(const Chirp& c, const ScanSet& set, int C) { return td_scan_layout(c, set, C).total; }

hipError_t launch_td_scan(const Chirp& c, const ScanSet& set, const float* ctr, const float* nrm, int C, int chunk,
                          void* ws, float* out, hipStream_t st) {
    const TdScanSet t = td_scan_set(set, chunk);
    const TdScanWorkspace lay = td_scan_layout(c, set, C);
    char* base = static_cast<char*>(ws);
    float* partial = reinterpret_cast<float*>(base + lay.partial);
    float2* U = reinterpret_cast<float2*>(base + lay.u);
    uint4* ops = reinterpret_cast<uint4*>(base + lay.ops);
    float* scales = reinterpret_cast<float*>(base + lay.scales);
    const size_t dsmem = sizeof(float2) * (c.N + 16 * (size_t)(c.K + 1));
    hipError_t e = allow_lds((const void*)dft_rows_kernel, dsmem);
    if (e != hipSuccess) return e;
    for (int i = 0; i < set.count; ++i) {
        const int A = set.apertures[i];
        hipLaunchKernelGGL(dft_rows_kernel, dim3((A + 15) / 16), dim3(256), dsmem, st, set.residuals[i], A, c.K, c.N,
                           1.0f, 1.0f, c.S, U + (size_t)t.ap_off[i] * c.S);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int apertures = t.ap_off[set.count];
    const int nrange = t.range_first[set.count];
    const dim3 grid((C + kScanCandidates - 1) / kScanCandidates, nrange);
    // the G x T operand tiles of a group: 16 x 128 uint4 for either sample count
    const size_t smem = sizeof(float2) * c.N + sizeof(float) * 16 + sizeof(uint4) * (size_t)kTdVjpGroup * 128;
    const bool fres = c.fresnel_n > 1.0f;
    if (c.S == 128) {
        hipLaunchKernelGGL(scan_operands_kernel<8>, dim3(apertures), dim3(64), 0, st, U, c.S, c.N, c.K, ops, scales);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        const void* k = fres ? (const void*)scan_td_kernel<8, true> : (const void*)scan_td_kernel<8, false>;
        if ((e = allow_lds(k, smem)) != hipSuccess) return e;
        if (fres)
            hipLaunchKernelGGL((scan_td_kernel<8, true>), grid, dim3(256), smem, st, c, t, ctr, nrm, C, ops, scales,
                               partial);
        else
            hipLaunchKernelGGL((scan_td_kernel<8, false>), grid, dim3(256), smem, st, c, t, ctr, nrm, C, ops, scales,
                               partial);
    } else {
        hipLaunchKernelGGL(scan_operands_kernel<16>, dim3(apertures), dim3(128), 0, st, U, c.S, c.N, c.K, ops, scales);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        const void* k = fres ? (const void*)scan_td_kernel<16, true> : (const void*)scan_td_kernel<16, false>;
        if ((e = allow_lds(k, smem)) != hipSuccess) return e;
        if (fres)
            hipLaunchKernelGGL((scan_td_kernel<16, true>), grid, dim3(256), smem, st, c, t, ctr, nrm, C, ops, scales,
                               partial);
        else
            hipLaunchKernelGGL((scan_td_kernel<16, false>), grid, dim3(256), smem, st, c, t, ctr, nrm, C, ops, scales,
                               partial);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_td_reduce, dim3((C + 3) / 4, set.count), dim3(256), 0, st, partial, C, nrange, t, out);
    return hipGetLastError();
}

}  // namespace dsplat

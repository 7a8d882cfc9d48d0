// Particle prep (edge mean, pad, phase origin, r2c FFT, CTF sign, scale) and
// finish (weight normalisation, 3-D c2r FFT, depad, phase origin) kernels.
//
// Prep follows programs/e2make3dpar.py:630-654,704-716 and
// FourierReconstructor::preprocess_slice (libEM/reconstructor.cpp:1022-1061);
// finish follows libEM/reconstructor.cpp:315-388,1552-1641.  All FFTs are
// LDS Stockham transforms (fh_device.hpp) over tiles of T transforms.
#include "fh_device.hpp"

namespace fh {

// ---------------------------------------------------------------- plans
bool make_plan(int N, FftPlan *p)
{
    p->N = N;
    p->npass = 0;
    int m = N;
    while (m % 8 == 0 && p->npass < 16) { p->radix[p->npass++] = 8; m /= 8; }
    while (m % 6 == 0 && p->npass < 16) { p->radix[p->npass++] = 6; m /= 6; }
    while (m % 4 == 0 && p->npass < 16) { p->radix[p->npass++] = 4; m /= 4; }
    while (m % 2 == 0 && p->npass < 16) { p->radix[p->npass++] = 2; m /= 2; }
    while (m % 3 == 0 && p->npass < 16) { p->radix[p->npass++] = 3; m /= 3; }
    while (m % 5 == 0 && p->npass < 16) { p->radix[p->npass++] = 5; m /= 5; }
    while (m % 7 == 0 && p->npass < 16) { p->radix[p->npass++] = 7; m /= 7; }
    while (m % 11 == 0 && p->npass < 16) { p->radix[p->npass++] = 11; m /= 11; }
    while (m % 13 == 0 && p->npass < 16) { p->radix[p->npass++] = 13; m /= 13; }
    if (m != 1 || N <= 1 || N >= 65536) return false;
    int ns = 1;
    for (int i = 0; i < p->npass; i++) {
        p->ns[i] = ns;
        p->mag[i] = ns == 1 ? 0u : (unsigned)((((uint64_t)1 << 32) + ns - 1) / ns);
        p->twm[i] = N / (ns * p->radix[i]);
        ns *= p->radix[i];
    }
    return true;
}

// Transforms per block: 4 for the row pass and the finish passes (the
// in-place tile, 5 P float2 with the twiddles, lets 8 blocks share a CU),
// 8 for the column pass up to P = 400 (its CTF epilogue keeps it busier per
// block).  Pads above 3200 / T take the ping-pong form.
static int pick_T(int P)
{
    (void)P;
    return 4;
}
static int pick_T_cols(int P) { return P <= 400 ? 8 : 4; }

// twiddles + one tile (in place) or two (ping-pong), see fft_tile
static size_t fft_lds_bytes(int P, int T)
{
    return sizeof(float2) * (size_t)P * (P * T <= 3200 ? T + 1 : 2 * T + 1);
}

// ---------------------------------------------------------------- edge mean
// EMData::get_edge_mean, 2-D branch (libEM/emdata_metadata.cpp:713-736):
// float pair sums of opposite edge pixels accumulated in double.
__global__ void __launch_bounds__(256) k_edge_mean(const float *__restrict__ imgs, int B, float *__restrict__ means)
{
    const float *d = imgs + (size_t)blockIdx.x * B * B;
    double acc = 0.0;
    for (int i = threadIdx.x; i < 2 * B; i += blockDim.x) {
        float pair;
        if (i < B) pair = d[i] + d[(size_t)(B - 1) * B + i];
        else {
            int r = i - B;
            pair = d[(size_t)r * B] + d[(size_t)r * B + B - 1];
        }
        acc += (double)pair;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) means[blockIdx.x] = (float)red[0] / (float)(B * 2 + B * 2);
}

hipError_t launch_edge_mean(const float *imgs, int n, int B, float *means, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_edge_mean, dim3(n), dim3(256), 0, s, imgs, B, means);
    return hipGetLastError();
}

// ---------------------------------------------------------------- row pass
// Each block transforms 2T padded rows of one particle: the rows are packed
// pairwise into T complex transforms z = a + i c (two-for-one real FFT).
// Input element xs of padded row yp is the phase-origin shifted
// (ifftshift) padded image; output row ys = (yp + P/2) mod P.
// NF: the pad as a compile-time constant (0: the runtime plan and P).
template <int T, int NF>
__global__ void __launch_bounds__(256) k_prep_rows(const float *__restrict__ imgs, int B, int Pr, int off,
                                                   const float *__restrict__ means, float2 *__restrict__ slices,
                                                   FftPlan plan, const float2 *__restrict__ twg)
{
    const int P = NF ? NF : Pr;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    float2 *Bf = A + T * P;
    const int H = P / 2;
    const int p = blockIdx.y;
    const int b0 = blockIdx.x * 2 * T;
    load_twiddles(tw, twg, P);
    const float mean = means[p];
    const float *img = imgs + (size_t)p * B * B;
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int t = idx % T;  // transform fastest: consecutive LDS slots
        const int xs = idx / T;
        const int xp = xs + H < P ? xs + H : xs + H - P;
        const int bx = xp - off;
        float re = 0.f, im = 0.f;
        if (bx >= 0 && bx < B) {
            const int r0 = b0 + 2 * t;
            if (r0 < B) re = img[(size_t)r0 * B + bx] - mean;
            if (r0 + 1 < B) im = img[(size_t)(r0 + 1) * B + bx] - mean;
        }
        A[xs * T + t] = make_float2(re, im);
    }
    __syncthreads();
    float2 *R = A;
    if constexpr (NF != 0) fft_fixed<T, NF>(A, tw);
    else R = fft_tile<T>(A, Bf, plan, tw);
    float2 *out = slices + (size_t)p * P * (H + 1);
    for (int idx = threadIdx.x; idx < T * (H + 1); idx += blockDim.x) {
        const int t = idx % T;
        const int k = idx / T;
        const int r0 = b0 + 2 * t;
        if (r0 >= B) continue;
        const float2 z1 = R[k * T + t];
        const float2 z2 = R[(k == 0 ? 0 : P - k) * T + t];
        const float2 a = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
        const float2 c = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
        int ys0 = r0 + off + H;  // < 2P
        if (ys0 >= P) ys0 -= P;
        out[(size_t)ys0 * (H + 1) + k] = a;
        if (r0 + 1 < B) {
            const int ys1 = ys0 + 1 == P ? 0 : ys0 + 1;
            out[(size_t)ys1 * (H + 1) + k] = c;
        }
    }
}

// Persistent row pass for the compile-time plans: each block walks tiles
// (particle, 2T rows) and reads the next tile's pixels into registers while
// it transforms the current one, so the global read latency hides behind
// the FFT.  Same arithmetic as k_prep_rows.
template <int T, int NF>
__global__ void __launch_bounds__(256) k_prep_rows_pipe(const float *__restrict__ imgs, int B, int off, int n,
                                                        const float *__restrict__ means, float2 *__restrict__ slices,
                                                        const float2 *__restrict__ twg)
{
    constexpr int P = NF, H = P / 2, KX = (P + 255) / 256;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    load_twiddles(tw, twg, P);
    const int nrb = (B + 2 * T - 1) / (2 * T);
    const int ntiles = nrb * n;
    // raw pixels of a tile: thread x-slot k holds padded column xs = tid + 256 k of all 2T rows, so
    // each load instruction reads 64 consecutive pixels of one row; out-of-image slots hold the mean
    // (zero after the subtraction)
    float v[KX][2 * T];
    float vmean = 0.f;
    auto fetch = [&](int tile) {
        const int p = tile / nrb, b0 = (tile - p * nrb) * 2 * T;
        const float mean = means[p];
        const float *img = imgs + (size_t)p * B * B;
        vmean = mean;
#pragma unroll
        for (int k = 0; k < KX; k++) {
            const int xs = threadIdx.x + 256 * k;
            const int xp = xs + H < P ? xs + H : xs + H - P;
            const int bx = xp - off;
            const bool in = xs < P && bx >= 0 && bx < B;
#pragma unroll
            for (int e = 0; e < 2 * T; e++) v[k][e] = in && b0 + e < B ? img[(size_t)(b0 + e) * B + bx] : mean;
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int p = tile / nrb, b0 = (tile - p * nrb) * 2 * T;
        const float mean = vmean;
#pragma unroll
        for (int k = 0; k < KX; k++) {
            const int xs = threadIdx.x + 256 * k;
            if (xs < P) {
                // transform t takes rows b0 + 2t (real part) and b0 + 2t + 1 (imaginary part)
                float4 *dst = reinterpret_cast<float4 *>(A + xs * T);
#pragma unroll
                for (int t2 = 0; t2 < T / 2; t2++)
                    dst[t2] = make_float4(v[k][4 * t2] - mean, v[k][4 * t2 + 1] - mean, v[k][4 * t2 + 2] - mean,
                                          v[k][4 * t2 + 3] - mean);
            }
        }
        __syncthreads();
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        fft_fixed<T, NF>(A, tw);
        float2 *out = slices + (size_t)p * P * (H + 1);
        for (int idx = threadIdx.x; idx < T * (H + 1); idx += 256) {
            const int t = idx % T;
            const int k = idx / T;
            const int r0 = b0 + 2 * t;
            if (r0 >= B) continue;
            const float2 z1 = A[k * T + t];
            const float2 z2 = A[(k == 0 ? 0 : P - k) * T + t];
            const float2 a = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
            const float2 c = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
            int ys0 = r0 + off + H;  // < 2P
            if (ys0 >= P) ys0 -= P;
            out[(size_t)ys0 * (H + 1) + k] = a;
            if (r0 + 1 < B) {
                const int ys1 = ys0 + 1 == P ? 0 : ys0 + 1;
                out[(size_t)ys1 * (H + 1) + k] = c;
            }
        }
        __syncthreads();  // the tile is read before the next one is stored
    }
}

// Compute units of the current device (persistent grids), cached per device.
static int device_cus()
{
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cus[dev] = v;
    }
    return cus[dev];
}

// Pads with a compile-time plan; others take the runtime plan.
#define FH_FIXED_PADS(X) X(4, 96) X(4, 192) X(4, 288) X(4, 384) X(4, 480) X(4, 600)
#define FH_FIXED_PADS_COLS(X) X(8, 96) X(8, 192) X(8, 288) X(8, 384) X(4, 480) X(4, 600)

hipError_t launch_prep_rows(const float *imgs, int n, int B, int P, const float *means, float2 *slices,
                            const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    // centre zero-pad (clip_inplace Region((B-P)//2, ...), e2make3dpar.py:651-652)
    const int d = B - P;
    const int fl = d >= 0 ? d / 2 : -((-d + 1) / 2);
    const int off = -fl;
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    dim3 grid((B + 2 * T - 1) / (2 * T), n);
    const int ntiles = (int)grid.x * n;
#define FH_ROWS_CASE(TT, NN)                                                                                    \
    if (T == TT && P == NN) {                                                                                   \
        int per = 0;                                                                                            \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_prep_rows_pipe<TT, NN>, 256, lds) != hipSuccess || \
            per < 1)                                                                                            \
            per = 1;                                                                                            \
        const int pgrid = std::min(ntiles, device_cus() * per);                                                 \
        hipLaunchKernelGGL((k_prep_rows_pipe<TT, NN>), dim3(pgrid), dim3(256), lds, s, imgs, B, off, n, means, slices, \
                           tw);                                                                                 \
        return hipGetLastError();                                                                               \
    }
    FH_FIXED_PADS(FH_ROWS_CASE)
#undef FH_ROWS_CASE
    if (T == 8)
        hipLaunchKernelGGL((k_prep_rows<8, 0>), grid, dim3(256), lds, s, imgs, B, P, off, means, slices, plan, tw);
    else
        hipLaunchKernelGGL((k_prep_rows<4, 0>), grid, dim3(256), lds, s, imgs, B, P, off, means, slices, plan, tw);
    return hipGetLastError();
}

// ---------------------------------------------------------------- column pass
// CTF_SIGN of EMAN2Ctf::compute_2d_complex (libEM/ctf.cpp:1076-1090) at
// the Fourier pixel of radius hyp = Util::hypot_fast(x, y).  Evaluated with the reference's float
// operation order (no contraction); the sign of cos is taken in double.
// cos(t) < 0 exactly when floor(t / pi - 1/2) is even; away from the
// zeros of cos (|frac| > 1e-9, far above the ~1e-13 error of u) the parity
// decides, next to them cos itself is evaluated.  The single-precision
// u_f = fma(t, fl(1/pi), -1/2) is within 2 |u| 2^-24 < 5e-4 of u for
// |u| < 4096, so a fraction of u_f more than 1e-3 from an integer already
// fixes floor(u); only the rest take the double-precision test.
__device__ __forceinline__ float ctf_sign(float hyp, float defocus, const CtfConst &cc)
{
#pragma clang fp contract(off)
    const float s = hyp * cc.ds;
    const float gam = -cc.g1 * s * s * s * s + cc.g2 * defocus * s * s;
    const float tf = gam - cc.acac;
    const float uf = __fmaf_rn(tf, 0.318309886f, -0.5f);
    const float flf = floorf(uf);
    const float frf = uf - flf;
    if (fabsf(uf) < 4096.f && frf > 1e-3f && frf < 1.f - 1e-3f) return (((int)flf) & 1) == 0 ? -1.0f : 1.0f;
    const double t = (double)tf;
    const double u = t * 0.31830988618379067154 - 0.5;
    const double fl = floor(u);
    const double fr = u - fl;
    if (fr > 1e-9 && fr < 1.0 - 1e-9) return (((long long)fl) & 1) == 0 ? -1.0f : 1.0f;
    return cos(t) < 0 ? -1.0f : 1.0f;
}

// defocus2 (optional): a second defocus per particle; the flip is then the
// product of the two signs (the dfdf flip of e2make3dpar.py:698-709)
template <int T, int NF>
__global__ void __launch_bounds__(256) k_prep_cols(int B, int Pr, int off, float2 *__restrict__ slices,
                                                   const float *__restrict__ defocus,
                                                   const float *__restrict__ defocus2, CtfConst cc, float scale,
                                                   FftPlan plan, const float2 *__restrict__ twg)
{
    const int P = NF ? NF : Pr;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    float2 *Bf = A + T * P;
    const int H = P / 2;
    // XCD-aware order: the hardware deals consecutive blocks round-robin to
    // the 8 XCDs (separate L2s); blocks k, k+8, k+16, ... of one XCD take
    // consecutive (column group, particle) tiles, so the column groups that
    // share a cache line meet in the same L2.
    unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned per = (gridDim.x * gridDim.y) / 8;
    if (lin < 8 * per) lin = (lin % 8) * per + lin / 8;
    const int p = (int)(lin / gridDim.x);
    const int x0 = (int)(lin % gridDim.x) * T;
    load_twiddles(tw, twg, P);
    float2 *sl = slices + (size_t)p * P * (H + 1);
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int y = idx / T;
        const int t = idx - y * T;
        const int x = x0 + t;
        int b = y - off - H;  // in [-P, P/2)
        if (b < 0) b += P;
        float2 v = make_float2(0.f, 0.f);
        if (x <= H && b < B) v = sl[(size_t)y * (H + 1) + x];
        A[y * T + t] = v;
    }
    __syncthreads();
    float2 *R = A;
    if constexpr (NF != 0) fft_fixed<T, NF>(A, tw);
    else R = fft_tile<T>(A, Bf, plan, tw);
    const float df = defocus ? defocus[p] : 0.f;
    const float df2 = defocus2 ? defocus2[p] : 0.f;
    // the sign depends on x^2 + y^2 only: rows y and P - y (signed y and -y,
    // 0 < y < P/2) share one evaluation
    for (int idx = threadIdx.x; idx < T * (H + 1); idx += blockDim.x) {
        const int y = idx / T;  // 0 .. P/2; row P/2 is signed -P/2
        const int t = idx - y * T;
        const int x = x0 + t;
        if (x > H) continue;
        // (v * (sg + 0i)) * scale with sg = +-1 equals v * (sg * scale) exactly
        // (up to the sign of a zero)
        const float hyp = defocus ? sqrt_int_rn(x * x + y * y) : 0.f;  // Util::hypot_fast(x, y)
        float ss = defocus ? ctf_sign(hyp, df, cc) * scale : scale;
        if (defocus2) ss = ss * ctf_sign(hyp, df2, cc);
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int row = m == 0 ? y : P - y;
            if (m == 1 && (y == 0 || y == H)) break;
            const float2 v = R[row * T + t];
            sl[(size_t)row * (H + 1) + x] = make_float2(v.x * ss, v.y * ss);
        }
    }
}

hipError_t launch_prep_cols(int n, int B, int P, float2 *slices, const float *defocus, const float *defocus2,
                            CtfConst cc, float scale, const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int d = B - P;
    const int fl = d >= 0 ? d / 2 : -((-d + 1) / 2);
    const int off = -fl;
    const int T = pick_T_cols(P);
    const int H = P / 2;
    const size_t lds = fft_lds_bytes(P, T);
    dim3 grid((H + 1 + T - 1) / T, n);
#define FH_COLS_CASE(TT, NN)                                                                                      \
    if (T == TT && P == NN) {                                                                                     \
        hipLaunchKernelGGL((k_prep_cols<TT, NN>), grid, dim3(256), lds, s, B, P, off, slices, defocus, defocus2, cc, \
                           scale, plan, tw);                                                                      \
        return hipGetLastError();                                                                                 \
    }
    FH_FIXED_PADS_COLS(FH_COLS_CASE)
#undef FH_COLS_CASE
    if (T == 8)
        hipLaunchKernelGGL((k_prep_cols<8, 0>), grid, dim3(256), lds, s, B, P, off, slices, defocus, defocus2, cc, scale,
                           plan, tw);
    else
        hipLaunchKernelGGL((k_prep_cols<4, 0>), grid, dim3(256), lds, s, B, P, off, slices, defocus, defocus2, cc, scale,
                           plan, tw);
    return hipGetLastError();
}

// ---------------------------------------------------------------- column-first prep
// The compile-time pads take the 2-D r2c transform columns first, so that
// the larger second pass reads and writes whole rows.
// Pass 1 (k_prep_ycols): 2T image columns per block, packed pairwise into T
// complex transforms along y (padded row ys = (b + off + P/2) mod P of image
// row b, edge mean subtracted, zero outside the image); the half spectrum
// ky = 0..P/2 of each real column is split out of the pair and stored to the
// intermediate I[p][ky][c] (row pitch IP, c = image column), 2T columns of a
// row per store run.
// Pass 2 (k_prep_xrows): T intermediate rows per block, transformed along x
// (padded column xs = (c + off + P/2) mod P); slice row ky takes
// kx = 0..P/2 and, for 0 < ky < P/2, row P - ky takes the conjugates of
// kx' = (P - kx) mod P (F(kx, -ky) = conj F(-kx, ky) for a real image), both
// times the CTF sign and scale of (kx, |ky|), evaluated once for the pair.
// The same output as k_prep_rows + k_prep_cols up to float rounding.
// Persistent blocks: the tiles of XCD x (blocks b = x mod 8) are the
// contiguous range [x ntiles / 8, (x + 1) ntiles / 8), so the tiles of a
// particle that share cache lines meet in the same L2; each block reads the
// next tile into registers while it transforms and stores the current one.
struct TileWalk {
    int t, end, step;
    __device__ TileWalk(int ntiles)
    {
        const int G8 = gridDim.x / 8, x = blockIdx.x % 8, lb = blockIdx.x / 8;
        const int lo = (int)((long long)ntiles * x / 8), hi = (int)((long long)ntiles * (x + 1) / 8);
        t = lo + lb;
        end = hi;
        step = G8;
    }
};

template <int T, int NF>
__global__ void __launch_bounds__(256) k_prep_ycols(const float *__restrict__ imgs, int B, int off, int n,
                                                    const float *__restrict__ means, float2 *__restrict__ inter,
                                                    int IP, const float2 *__restrict__ twg)
{
    constexpr int P = NF, H = P / 2, E = (T * P + 255) / 256;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    load_twiddles(tw, twg, P);
    const int ng = (B + 2 * T - 1) / (2 * T);
    TileWalk w(ng * n);
    // element e of the thread: (ys, t) = ((tid + 256 e) / T, (tid + 256 e) % T), image columns c, c + 1
    // of padded row ys; slots outside the image hold the mean (zero after the subtraction)
    float2 v[E];
    float vmean = 0.f;
    auto fetch = [&](int tile) {
        const int p = tile / ng, c0 = (tile - p * ng) * 2 * T;
        const float mean = means[p];
        const float *img = imgs + (size_t)p * B * B;
        vmean = mean;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = threadIdx.x + 256 * e;
            const int t = idx % T, ys = idx / T;
            int b = ys - off - H;  // image row of padded row ys, in [-P, P/2)
            if (b < 0) b += P;
            const int c = c0 + 2 * t;
            const bool in = idx < T * P && b < B;
            v[e].x = in && c < B ? img[(size_t)b * B + c] : mean;
            v[e].y = in && c + 1 < B ? img[(size_t)b * B + c + 1] : mean;
        }
    };
    if (w.t < w.end) fetch(w.t);
    for (; w.t < w.end; w.t += w.step) {
        const int p = w.t / ng, c0 = (w.t - p * ng) * 2 * T;
        const float mean = vmean;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = threadIdx.x + 256 * e;
            if (E * 256 == T * P || idx < T * P) A[idx] = make_float2(v[e].x - mean, v[e].y - mean);  // A[ys T + t]
        }
        __syncthreads();
        if (w.t + w.step < w.end) fetch(w.t + w.step);
        fft_fixed<T, NF>(A, tw);
        float2 *I = inter + (size_t)p * (H + 1) * IP;
        for (int idx = threadIdx.x; idx < T * (H + 1); idx += 256) {
            const int t = idx % T;
            const int k = idx / T;
            const int c = c0 + 2 * t;
            if (c >= B) continue;
            const float2 z1 = A[k * T + t];
            const float2 z2 = A[(k == 0 ? 0 : P - k) * T + t];
            // columns c and c + 1 (for odd B the last pair's second column is
            // zero and lands in the row padding, IP > B)
            *reinterpret_cast<float4 *>(I + (size_t)k * IP + c) = make_float4(
                0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y), 0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
        }
        __syncthreads();  // the tile is read before the next one is stored
    }
}

template <int T, int NF>
__global__ void __launch_bounds__(256) k_prep_xrows(int B, int off, int n, const float2 *__restrict__ inter, int IP,
                                                    float2 *__restrict__ slices, const float *__restrict__ defocus,
                                                    const float *__restrict__ defocus2, CtfConst cc, float scale,
                                                    const float2 *__restrict__ twg)
{
    constexpr int P = NF, H = P / 2, E = (T * P + 255) / 256, NG = (H + 1 + T - 1) / T;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    load_twiddles(tw, twg, P);
    TileWalk w(NG * n);
    // element e of the thread: (xs, t) = ((tid + 256 e) / T, (tid + 256 e) % T): a wave reads runs of
    // 64 / T columns of T intermediate rows and writes LDS without bank conflicts
    float2 v[E];
    auto fetch = [&](int tile) {
        const int p = tile / NG, k0 = (tile - p * NG) * T;
        const float2 *I = inter + (size_t)p * (H + 1) * IP;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = threadIdx.x + 256 * e;
            const int t = idx % T, xs = idx / T;
            const int k = k0 + t;
            int c = xs - off - H;  // image column of padded column xs
            if (c < 0) c += P;
            v[e] = (idx < T * P && k <= H && c < B) ? I[(size_t)k * IP + c] : make_float2(0.f, 0.f);
        }
    };
    if (w.t < w.end) fetch(w.t);
    for (; w.t < w.end; w.t += w.step) {
        const int p = w.t / NG, k0 = (w.t - p * NG) * T;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = threadIdx.x + 256 * e;
            if (E * 256 == T * P || idx < T * P) A[idx] = v[e];  // A[xs T + t]
        }
        __syncthreads();
        if (w.t + w.step < w.end) fetch(w.t + w.step);
        fft_fixed<T, NF>(A, tw);
        const float df = defocus ? defocus[p] : 0.f;
        const float df2 = defocus2 ? defocus2[p] : 0.f;
        float2 *sl = slices + (size_t)p * P * (H + 1);
        // two rows per lane (one 16-byte LDS read per column), lanes along the
        // rows: whole-row store runs
        static_assert(T % 2 == 0, "row pairs");
        for (int idx = threadIdx.x; idx < (T / 2) * (H + 1); idx += 256) {
            const int t0 = 2 * (idx / (H + 1));
            const int kx = idx - (t0 / 2) * (H + 1);
            const float4 u2 = *reinterpret_cast<const float4 *>(A + kx * T + t0);
            const float4 q2 = *reinterpret_cast<const float4 *>(A + (kx == 0 ? 0 : P - kx) * T + t0);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int ky = k0 + t0 + h;
                if (ky > H) continue;
                const float hyp = defocus ? sqrt_int_rn(kx * kx + ky * ky) : 0.f;  // Util::hypot_fast(x, y)
                float ss = defocus ? ctf_sign(hyp, df, cc) * scale : scale;
                if (defocus2) ss = ss * ctf_sign(hyp, df2, cc);
                const float ux = h ? u2.z : u2.x, uy = h ? u2.w : u2.y;
                const float qx = h ? q2.z : q2.x, qy = h ? q2.w : q2.y;
                sl[(size_t)ky * (H + 1) + kx] = make_float2(ux * ss, uy * ss);
                if (ky > 0 && ky < H) sl[(size_t)(P - ky) * (H + 1) + kx] = make_float2(qx * ss, -qy * ss);
            }
        }
        __syncthreads();  // the tile is read before the next one is stored
    }
}

// Persistent grid of a column-first pass: resident blocks per CU times the
// CUs, a multiple of 8 (one share per XCD), at most the tiles rounded up to 8.
template <typename K>
static int persistent_grid(K kernel, size_t lds, int ntiles)
{
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, lds) != hipSuccess || per < 1) per = 1;
    const int g = std::min(device_cus() * per, (ntiles + 7) / 8 * 8);
    return std::max(8, g / 8 * 8);
}

// Intermediate row pitch of box B (a multiple of 16 columns: 128-byte rows,
// and one padding column for odd B).
int prep_inter_pitch(int B) { return (B + 16) & ~15; }

bool prep_column_first(int P)
{
#define FH_CF_CASE(TT, NN) if (P == NN) return true;
    FH_FIXED_PADS(FH_CF_CASE)
#undef FH_CF_CASE
    return false;
}

// The whole prep of n particles: column-first through `inter` (room for
// inter_n particles of box P) in sub-batches for the compile-time pads, else
// the row pass and the column pass in place.
hipError_t launch_prep2d(const float *imgs, int n, int B, int P, const float *means, float2 *slices,
                         float2 *inter, int inter_n, const float *defocus, const float *defocus2, CtfConst cc,
                         float scale, const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (!inter || inter_n <= 0 || !prep_column_first(P) || B > P) {
        hipError_t e = launch_prep_rows(imgs, n, B, P, means, slices, plan, tw, s);
        if (e != hipSuccess) return e;
        return launch_prep_cols(n, B, P, slices, defocus, defocus2, cc, scale, plan, tw, s);
    }
    const int d = B - P;
    const int fl = d >= 0 ? d / 2 : -((-d + 1) / 2);  // (B - P) // 2, as in the row pass
    const int off = -fl;
    const int H = P / 2, IP = prep_inter_pitch(B);
    // particles per sub-batch: inter holds inter_n particles at the pitch of box P
    const long long room = (long long)inter_n * prep_inter_pitch(P);
    const int sub = (int)std::max(1LL, room / IP);
    for (int i0 = 0; i0 < n; i0 += sub) {
        const int m = std::min(sub, n - i0);
        const float *im = imgs + (size_t)i0 * B * B;
        const float *mn = means + i0;
        float2 *sl = slices + (size_t)i0 * P * (H + 1);
        const float *df = defocus ? defocus + i0 : nullptr;
        const float *df2 = defocus2 ? defocus2 + i0 : nullptr;
#define FH_CF_LAUNCH(TT, NN)                                                                                      \
        if (P == NN) {                                                                                            \
            constexpr int T1 = 8, T2 = 4;                                                         \
            const size_t l1 = sizeof(float2) * (T1 + 1) * NN, l2 = sizeof(float2) * (T2 + 1) * NN;                \
            const int g1 = persistent_grid(k_prep_ycols<T1, NN>, l1, (B + 2 * T1 - 1) / (2 * T1) * m);              \
            const int g2 = persistent_grid(k_prep_xrows<T2, NN>, l2, (H + T2) / T2 * m);                          \
            hipLaunchKernelGGL((k_prep_ycols<T1, NN>), dim3(g1), dim3(256), l1, s, im, B, off, m, mn, inter, IP, tw); \
            hipLaunchKernelGGL((k_prep_xrows<T2, NN>), dim3(g2), dim3(256), l2, s, B, off, m, inter, IP, sl, df, df2, \
                               cc, scale, tw);                                                                    \
        }
        FH_FIXED_PADS(FH_CF_LAUNCH)
#undef FH_CF_LAUNCH
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------- normalise
// ReconstructorVolumeData::normalize_threed (libEM/reconstructor.cpp:315-388).
// The conjugate-plane loop pairs each weight on the x=0 and x=P/2 planes
// with its partner at (-y,-z); the pairing is an involution, so each fixed
// weight is norm[self] + norm[partner] except for self-paired cells, of which
// five are doubled.  Step 1 writes the fixed planes to plane_tmp [2][P][P].
__global__ void k_norm_planes(const float *__restrict__ norm, int P, float *__restrict__ plane_tmp)
{
    const int H = P / 2;
    const size_t nnx = H + 1;
    const int ys = blockIdx.x * blockDim.x + threadIdx.x;
    const int zs = blockIdx.y;
    const int pl = blockIdx.z;  // 0: x=0, 1: x=H
    if (ys >= P) return;
    const int xs = pl ? H : 0;
    const int yp = (P - ys) % P, zp = (P - zs) % P;
    const float self = norm[((size_t)zs * P + ys) * nnx + xs];
    float v;
    if (yp == ys && zp == zs) {
        bool dbl;
        if (pl == 0) dbl = (ys == 0 && zs == H) || (ys == H && zs == 0) || (ys == H && zs == H);
        else dbl = (ys == 0 && zs == H) || (ys == H && zs == 0);
        v = dbl ? self * 2 : self;
    } else {
        v = self + norm[((size_t)zp * P + yp) * nnx + xs];
    }
    plane_tmp[((size_t)pl * P + zs) * P + ys] = v;
}

__global__ void k_normalize(float2 *__restrict__ data, float *__restrict__ norm, int P, int sqrtnorm,
                            const float *__restrict__ plane_tmp)
{
    const int H = P / 2;
    const size_t nnx = H + 1;
    const size_t total = nnx * P * P;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int xs = (int)(i % nnx);
        const size_t yz = i / nnx;
        float d;
        if (xs == 0 || xs == H) {
            const int ys = (int)(yz % P), zs = (int)(yz / P);
            d = plane_tmp[((size_t)(xs == 0 ? 0 : 1) * P + zs) * P + ys];
            norm[i] = d;
        } else {
            d = norm[i];
        }
        if (sqrtnorm) d = sqrt_rn(d);
        float2 v = data[i];
        if (d == 0) v = make_float2(0.f, 0.f);
        else v = make_float2(__fdiv_rn(v.x, d), __fdiv_rn(v.y, d));
        data[i] = v;
    }
}

hipError_t launch_normalize(float2 *data, float *norm, int P, int sqrtnorm, float *plane_tmp, hipStream_t s)
{
    hipLaunchKernelGGL(k_norm_planes, dim3((P + 127) / 128, P, 2), dim3(128), 0, s, norm, P, plane_tmp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_normalize, dim3(4096), dim3(256), 0, s, data, norm, P, sqrtnorm, plane_tmp);
    return hipGetLastError();
}

// ---------------------------------------------------------------- 3-D inverse
// In-place inverse complex transform of length P along z (pass 0) or y
// (pass 1) for tiles of T consecutive x; inverse = conj o forward o conj.
// FWD: the forward transform instead (no conjugation), each output n times
// (-1)^n: the axis's phase origin moved to the corner
template <int T, int FWD = 0>
__global__ void __launch_bounds__(256) k_fft_axis_inv(float2 *__restrict__ data, int P, size_t axis_stride,
                                                      size_t outer_stride, FftPlan plan,
                                                      const float2 *__restrict__ twg)
{
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    float2 *Bf = A + T * P;
    const int H = P / 2;
    const int x0 = blockIdx.x * T;
    float2 *base = data + (size_t)blockIdx.y * outer_stride;
    load_twiddles(tw, twg, P);
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int n = idx / T;
        const int t = idx - n * T;
        const int x = x0 + t;
        float2 v = make_float2(0.f, 0.f);
        if (x <= H) v = FWD ? base[(size_t)n * axis_stride + x] : conjf2(base[(size_t)n * axis_stride + x]);
        A[n * T + t] = v;
    }
    __syncthreads();
    float2 *R = fft_tile<T>(A, Bf, plan, tw);
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int n = idx / T;
        const int t = idx - n * T;
        const int x = x0 + t;
        if (x <= H) {
            const float2 r = R[n * T + t];
            base[(size_t)n * axis_stride + x] = FWD ? ((n & 1) ? make_float2(-r.x, -r.y) : r) : conjf2(r);
        }
    }
}

hipError_t launch_fft3d_inverse_zy(float2 *data, int P, const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    const int H = P / 2;
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    const size_t row = H + 1;
    dim3 grid((H + 1 + T - 1) / T, P);
    // z pass: tiles at fixed y
    if (T == 8)
        hipLaunchKernelGGL(k_fft_axis_inv<8>, grid, dim3(256), lds, s, data, P, (size_t)P * row, row, plan, tw);
    else
        hipLaunchKernelGGL(k_fft_axis_inv<4>, grid, dim3(256), lds, s, data, P, (size_t)P * row, row, plan, tw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // y pass: tiles at fixed z
    if (T == 8)
        hipLaunchKernelGGL(k_fft_axis_inv<8>, grid, dim3(256), lds, s, data, P, row, (size_t)P * row, plan, tw);
    else
        hipLaunchKernelGGL(k_fft_axis_inv<4>, grid, dim3(256), lds, s, data, P, row, (size_t)P * row, plan, tw);
    return hipGetLastError();
}

// c2r along x for row pairs, x 1/P^3 (do_ift_inplace, emdata_transform.cpp:392-425),
// depad, and phase origin to centre (fftshift) on the write.
// nz planes of P rows; shift 1: fftshift of the written volume (3-D finish),
// shift 2: fftshift of each plane (a batch of 2-D images), 0: rows written
// in place (2-D projection).
template <int T>
__global__ void __launch_bounds__(256) k_c2r_x(const float2 *__restrict__ data, int P, int nz, int shift,
                                               float *__restrict__ out, float scale, FftPlan plan,
                                               const float2 *__restrict__ twg)
{
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + P;
    float2 *Bf = A + T * P;
    const int H = P / 2;
    const size_t nrows = (size_t)nz * P;
    const size_t row0 = (size_t)blockIdx.x * 2 * T;
    load_twiddles(tw, twg, P);
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int t = idx / P;
        const int k = idx - t * P;
        const size_t ra = row0 + 2 * t, rc = ra + 1;
        float2 a = make_float2(0.f, 0.f), c = make_float2(0.f, 0.f);
        if (ra < nrows) {
            a = k <= H ? data[ra * (H + 1) + k] : conjf2(data[ra * (H + 1) + (P - k)]);
            if (k == 0 || k == H) a.y = 0.f;
        }
        if (rc < nrows) {
            c = k <= H ? data[rc * (H + 1) + k] : conjf2(data[rc * (H + 1) + (P - k)]);
            if (k == 0 || k == H) c.y = 0.f;
        }
        // Z = A + iC, inverse via conj o forward o conj
        A[k * T + t] = make_float2(a.x - c.y, -(a.y + c.x));
    }
    __syncthreads();
    float2 *R = fft_tile<T>(A, Bf, plan, tw);
    for (int idx = threadIdx.x; idx < T * P; idx += blockDim.x) {
        const int t = idx / P;
        const int x = idx - t * P;
        const size_t ra = row0 + 2 * t;
        const float2 z = conjf2(R[x * T + t]);
        const int xc = x + H < P ? x + H : x + H - P;
        for (int q = 0; q < 2; q++) {
            const size_t r = ra + q;
            if (r >= nrows) break;
            if (!shift) {
                out[r * P + x] = (q == 0 ? z.x : z.y) * scale;
                continue;
            }
            const int y = (int)(r % P), zz = (int)(r / P);
            const int yc = y + H < P ? y + H : y + H - P;
            const int zc = shift == 2 ? zz : (zz + H < P ? zz + H : zz + H - P);
            out[((size_t)zc * P + yc) * P + xc] = (q == 0 ? z.x : z.y) * scale;
        }
    }
}

static hipError_t c2r_rows(const float2 *data, int P, int nz, int shift, float *out, float scale,
                           const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    const size_t nrows = (size_t)nz * P;
    dim3 grid((unsigned)((nrows + 2 * T - 1) / (2 * T)));
    if (T == 8)
        hipLaunchKernelGGL(k_c2r_x<8>, grid, dim3(256), lds, s, data, P, nz, shift, out, scale, plan, tw);
    else
        hipLaunchKernelGGL(k_c2r_x<4>, grid, dim3(256), lds, s, data, P, nz, shift, out, scale, plan, tw);
    return hipGetLastError();
}

hipError_t launch_c2r_x(const float2 *data, int P, float *out, float scale, const FftPlan &plan, const float2 *tw,
                        hipStream_t s)
{
    return c2r_rows(data, P, P, 1, out, scale, plan, tw, s);
}

// 2-D do_ift of one (P+2) x P complex image (emdata_transform.cpp:330-375):
// inverse FFT along y of the P/2+1 columns, c2r along x, x 1/P^2.
hipError_t launch_ift2d(float2 *data, int P, float *out, const FftPlan &plan, const float2 *tw, hipStream_t s)
{
    const int H = P / 2;
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    const size_t row = H + 1;
    dim3 grid((H + 1 + T - 1) / T, 1);
    if (T == 8)
        hipLaunchKernelGGL(k_fft_axis_inv<8>, grid, dim3(256), lds, s, data, P, row, (size_t)0, plan, tw);
    else
        hipLaunchKernelGGL(k_fft_axis_inv<4>, grid, dim3(256), lds, s, data, P, row, (size_t)0, plan, tw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return c2r_rows(data, P, 1, 0, out, 1.0f / ((float)P * (float)P), plan, tw, s);
}

// do_ift of n complex images (P x (P/2+1) each, in place: overwritten) into n
// real P x P images with scale, phase origin to the centre: the inverse of the
// row and column passes (the particle's flip back to real space)
hipError_t launch_ift2d_batch(float2 *data, int n, int P, float *out, float scale, const FftPlan &plan,
                              const float2 *tw, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int H = P / 2;
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    const size_t row = H + 1;
    dim3 grid((H + 1 + T - 1) / T, n);
    if (T == 8)
        hipLaunchKernelGGL(k_fft_axis_inv<8>, grid, dim3(256), lds, s, data, P, row, (size_t)P * row, plan, tw);
    else
        hipLaunchKernelGGL(k_fft_axis_inv<4>, grid, dim3(256), lds, s, data, P, row, (size_t)P * row, plan, tw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return c2r_rows(data, P, n, 2, out, scale, plan, tw, s);
}

// ---------------------------------------------------------------- forward 3-D
// do_fft of a real P^3 volume with the phase origin moved to the corner (the
// Fourier-space xform.phaseorigin.tocorner, center_origin_fft): the row and
// column passes of the particle prep over the P z planes (their load index
// applies the x and y shift), then the z transform, each output times
// (-1)^z.  vol: P^3 floats, out: P x P x (P/2+1) complex.
hipError_t launch_fft3d_forward_corner(const float *vol, int P, const float *zeros, float2 *out, const FftPlan &plan,
                                       const float2 *tw, hipStream_t s)
{
    hipError_t e = launch_prep_rows(vol, P, P, P, zeros, out, plan, tw, s);
    if (e != hipSuccess) return e;
    e = launch_prep_cols(P, P, P, out, nullptr, nullptr, CtfConst{}, 1.0f, plan, tw, s);
    if (e != hipSuccess) return e;
    const int H = P / 2;
    const int T = pick_T(P);
    const size_t lds = fft_lds_bytes(P, T);
    const size_t row = H + 1;
    dim3 grid((H + 1 + T - 1) / T, P);
    if (T == 8)
        hipLaunchKernelGGL((k_fft_axis_inv<8, 1>), grid, dim3(256), lds, s, out, P, (size_t)P * row, row, plan, tw);
    else
        hipLaunchKernelGGL((k_fft_axis_inv<4, 1>), grid, dim3(256), lds, s, out, P, (size_t)P * row, row, plan, tw);
    return hipGetLastError();
}

// ---------------------------------------------------------------- task merge
// e2spa_make3d.py:180-184: a task's normalised volume times its normout
// ("math.multamplitude": both floats of a voxel times its weight), added to
// the sum; the normout added to the summed weights
__global__ void k_merge_add(float2 *__restrict__ data, float *__restrict__ norm, const float2 *__restrict__ threed,
                            const float *__restrict__ normv, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float w = normv[i];
        const float2 t = threed[i];
        const float2 d = data[i];
        data[i] = make_float2(d.x + t.x * w, d.y + t.y * w);
        norm[i] = norm[i] + w;
    }
}

// e2spa_make3d.py:186-187: the summed weights' reciprocal ("math.reciprocal",
// 0 where the weight is 0), times the summed volume
__global__ void k_merge_scale(float2 *__restrict__ data, const float *__restrict__ norm, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float w = norm[i];
        const float r = w == 0.0f ? 0.0f : 1.0f / w;
        const float2 d = data[i];
        data[i] = make_float2(d.x * r, d.y * r);
    }
}

hipError_t launch_merge_add(float2 *data, float *norm, const float2 *threed, const float *normv, size_t n,
                            hipStream_t s)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_merge_add, dim3(blocks), dim3(256), 0, s, data, norm, threed, normv, n);
    return hipGetLastError();
}

hipError_t launch_merge_scale(float2 *data, const float *norm, size_t n, hipStream_t s)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_merge_scale, dim3(blocks), dim3(256), 0, s, data, norm, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- fills
// Fill kernels instead of hipMemsetAsync so the sequences stay capturable.
__global__ void k_fill_f32(float *p, float v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int *p, int v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_scale_f32(float *p, float v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = p[i] * v;
}

static unsigned fill_grid(size_t n)
{
    size_t g = (n + 255) / 256;
    return (unsigned)(g < 8192 ? (g ? g : 1) : 8192);
}

hipError_t launch_fill_f32(float *p, float v, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_f32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_scale_f32(float *p, float v, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_f32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// setup_seed weights (libEM/reconstructor.cpp:870-904): seed weight
// everywhere, halved on the x=0 and x=P/2 planes except four cells.
__global__ void k_seed_weights(float *norm, int P, float w)
{
    const int H = P / 2;
    const size_t nnx = H + 1;
    const size_t total = nnx * P * P;
    const float half = (float)(w / 2.0);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int xs = (int)(i % nnx);
        const size_t yz = i / nnx;
        const int ys = (int)(yz % P), zs = (int)(yz / P);
        float v = w;
        if (xs == 0 || xs == H) {
            v = half;
            if ((ys == H && zs == 0) || (ys == 0 && zs == H)) v = w;
        }
        norm[i] = v;
    }
}

hipError_t launch_seed_weights(float *norm, int P, float w, hipStream_t s)
{
    hipLaunchKernelGGL(k_seed_weights, dim3(4096), dim3(256), 0, s, norm, P, w);
    return hipGetLastError();
}

}  // namespace fh

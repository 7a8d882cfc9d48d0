// Device helpers shared by the fourier_hip kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "fh_launch.hpp"

namespace fh {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Correctly rounded float sqrt: sqrt of a float never lies within 2^-53 of a
// float rounding midpoint, so the double sqrt rounded once to float is exact
// (the single-precision hardware sqrt is only accurate to 1 ulp).  Equals
// hypot((float)x, (float)y) of the reference for integer x^2 + y^2 < 2^24.
__device__ __forceinline__ float sqrt_rn(float v) { return (float)sqrt((double)v); }
// The same for an integer 0 <= n < 2^24 in single precision: the hardware
// sqrt s (within 1 ulp) is moved to a neighbour when sqrt(n) lies beyond
// the midpoint m between them.  With s = M 2^E (E <= -11 here), n - m^2 is
// an odd multiple of h^2 (h = ulp / 2) and s * s_next = m^2 - h^2 (for a
// power of two the same holds with a smaller h), so the exactly-signed fma
// residual n - s * s_next decides n against m^2; no midpoint is ever hit.
__device__ __forceinline__ float sqrt_int_rn(int n)
{
    const float x = (float)n;
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1), su = __int_as_float(__float_as_int(s) + 1);
    float r = __fmaf_rn(-sd, s, x) <= 0.f ? sd : s;
    r = __fmaf_rn(-su, s, x) > 0.f ? su : r;
    return n == 0 ? 0.f : r;
}
// multiply by -i
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// ---------------------------------------------------------------------------
// Radix-R forward DFT on registers (exp(-2 pi i / R) kernel).  The constants
// are cos/sin of 2 pi k / R.
// ---------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void dft(float2 *v);

template <>
__device__ __forceinline__ void dft<2>(float2 *v)
{
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(float2 *v)
{
    const float s1 = 0.86602540378443864676f;  // sin(2 pi/3)
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    float2 d = csub(v[1], v[2]);
    float2 t3 = make_float2(s1 * d.x, s1 * d.y);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, mul_mi(t3));
    v[2] = csub(t2, mul_mi(t3));
}

template <>
__device__ __forceinline__ void dft<4>(float2 *v)
{
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, mul_mi(b1));
    v[3] = csub(a1, mul_mi(b1));
}

template <>
__device__ __forceinline__ void dft<5>(float2 *v)
{
    const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
    const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
    const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
    const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
    float2 a = v[0];
    float2 p14 = cadd(v[1], v[4]), m14 = csub(v[1], v[4]);
    float2 p23 = cadd(v[2], v[3]), m23 = csub(v[2], v[3]);
    float2 r1 = make_float2(a.x + c1 * p14.x + c2 * p23.x, a.y + c1 * p14.y + c2 * p23.y);
    float2 r2 = make_float2(a.x + c2 * p14.x + c1 * p23.x, a.y + c2 * p14.y + c1 * p23.y);
    float2 i1 = make_float2(s1 * m14.x + s2 * m23.x, s1 * m14.y + s2 * m23.y);
    float2 i2 = make_float2(s2 * m14.x - s1 * m23.x, s2 * m14.y - s1 * m23.y);
    v[0] = make_float2(a.x + p14.x + p23.x, a.y + p14.y + p23.y);
    v[1] = cadd(r1, mul_mi(i1));
    v[4] = csub(r1, mul_mi(i1));
    v[2] = cadd(r2, mul_mi(i2));
    v[3] = csub(r2, mul_mi(i2));
}

// Odd prime radices 7, 11, 13 (the remaining factors of EMAN2's good box
// sizes): X[0] = sum v; for k = 1..(R-1)/2, with p_j = v_j + v_{R-j} and
// m_j = v_j - v_{R-j}: X[k] = v_0 + sum_j cos(2 pi jk/R) p_j - i sum_j
// sin(2 pi jk/R) m_j and X[R-k] the same with +i.
constexpr float prime_cos(int R, int k)
{
    return R == 7 ? (k == 1 ? 0.6234898018587336f : k == 2 ? -0.22252093395631434f : -0.900968867902419f)
         : R == 11 ? (k == 1 ? 0.8412535328311812f : k == 2 ? 0.41541501300188644f : k == 3 ? -0.142314838273285f
                     : k == 4 ? -0.654860733945285f : -0.9594929736144974f)
         : (k == 1 ? 0.8854560256532099f : k == 2 ? 0.5680647467311559f : k == 3 ? 0.120536680255323f
            : k == 4 ? -0.35460488704253545f : k == 5 ? -0.7485107481711012f : -0.970941817426052f);
}
constexpr float prime_sin(int R, int k)
{
    return R == 7 ? (k == 1 ? 0.7818314824680298f : k == 2 ? 0.9749279121818236f : 0.43388373911755823f)
         : R == 11 ? (k == 1 ? 0.5406408174555976f : k == 2 ? 0.9096319953545183f : k == 3 ? 0.9898214418809328f
                     : k == 4 ? 0.7557495743542583f : 0.28173255684142967f)
         : (k == 1 ? 0.4647231720437685f : k == 2 ? 0.8229838658936564f : k == 3 ? 0.992708874098054f
            : k == 4 ? 0.9350162426854148f : k == 5 ? 0.6631226582407952f : 0.23931566428755768f);
}

template <int R>
__device__ __forceinline__ void dft_prime(float2 *v)
{
    constexpr int H = (R - 1) / 2;
    float2 p[H], m[H];
    const float2 x0 = v[0];
    float2 sum = v[0];
#pragma unroll
    for (int j = 0; j < H; j++) {
        p[j] = cadd(v[j + 1], v[R - 1 - j]);
        m[j] = csub(v[j + 1], v[R - 1 - j]);
        sum = cadd(sum, p[j]);
    }
#pragma unroll
    for (int k = 1; k <= H; k++) {
        float2 re = x0, im = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= H; j++) {
            const int a = (j * k) % R;  // cos/sin(2 pi a / R), folded into (0, R/2)
            const float cc = a <= H ? prime_cos(R, a) : prime_cos(R, R - a);
            const float ss = a <= H ? prime_sin(R, a) : -prime_sin(R, R - a);
            re.x += cc * p[j - 1].x;
            re.y += cc * p[j - 1].y;
            im.x += ss * m[j - 1].x;
            im.y += ss * m[j - 1].y;
        }
        v[k] = cadd(re, mul_mi(im));
        v[R - k] = csub(re, mul_mi(im));
    }
    v[0] = sum;
}
template <>
__device__ __forceinline__ void dft<7>(float2 *v) { dft_prime<7>(v); }
template <>
__device__ __forceinline__ void dft<11>(float2 *v) { dft_prime<11>(v); }
template <>
__device__ __forceinline__ void dft<13>(float2 *v) { dft_prime<13>(v); }

// 6 = 2 x 3: DFT3 of the even and odd halves, X[k] = E[k] + W6^k O[k],
// X[k+3] = E[k] - W6^k O[k].
template <>
__device__ __forceinline__ void dft<6>(float2 *v)
{
    const float h = 0.86602540378443864676f;  // sin(2 pi/6)
    float2 e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]};
    dft<3>(e);
    dft<3>(o);
    const float2 o1 = make_float2(0.5f * o[1].x + h * o[1].y, 0.5f * o[1].y - h * o[1].x);   // W6^1 = 1/2 - i h
    const float2 o2 = make_float2(-0.5f * o[2].x + h * o[2].y, -0.5f * o[2].y - h * o[2].x); // W6^2 = -1/2 - i h
    v[0] = cadd(e[0], o[0]);
    v[3] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[4] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[5] = csub(e[2], o2);
}

// 8 = 2 x 4: DFT4 of the even and odd halves, X[k] = E[k] + W8^k O[k],
// X[k+4] = E[k] - W8^k O[k].
template <>
__device__ __forceinline__ void dft<8>(float2 *v)
{
    const float r = 0.70710678118654752440f;  // 1/sqrt(2)
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft<4>(e);
    dft<4>(o);
    const float2 o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));    // W8^1 = (1 - i)/sqrt 2
    const float2 o2 = mul_mi(o[2]);                                                // W8^2 = -i
    const float2 o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));   // W8^3 = -(1 + i)/sqrt 2
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

// One Stockham autosort pass over T interleaved transforms of length N:
// element (t, n) lives at buf[n*T + t].  tw[m] = exp(-2 pi i m / N).  Pass
// p has stride Ns = prod(radix[<p]); j / Ns is taken with the plan's
// multiply-high constant.
template <int R, int T>
__device__ __forceinline__ void stockham_pass(const float2 *__restrict__ src, float2 *__restrict__ dst,
                                              int N, int Ns, unsigned mag, int twm, const float2 *__restrict__ tw)
{
    const int NR = N / R;
    const int total = NR * T;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int t = b % T;
        const int j = b / T;
        const int q = Ns == 1 ? j : (int)__umulhi((unsigned)j, mag);
        const int k = j - q * Ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            float2 a = src[(j + r * NR) * T + t];
            if (r > 0) a = cmul(a, tw[r * k * twm]);
            v[r] = a;
        }
        dft<R>(v);
        const int base = q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[(base + r * Ns) * T + t] = v[r];
    }
}

// In-place pass: every thread first reads and transforms all of its
// butterflies (at most G per thread, G * R <= 18 registers of float2), then,
// after a barrier, writes them back to their Stockham output positions in
// the same buffer.
template <int R, int T, int G>
__device__ __forceinline__ void stockham_pass_inplace(float2 *buf, int N, int Ns, unsigned mag, int twm,
                                                      const float2 *__restrict__ tw)
{
    const int NR = N / R;
    const int total = NR * T;
    float2 v[G][R];
    int dsti[G];
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
        const int b = threadIdx.x + gi * blockDim.x;
        dsti[gi] = -1;
        if (b < total) {
            const int t = b % T;
            const int j = b / T;
            const int q = Ns == 1 ? j : (int)__umulhi((unsigned)j, mag);
            const int k = j - q * Ns;
#pragma unroll
            for (int r = 0; r < R; r++) {
                float2 a = buf[(j + r * NR) * T + t];
                if (r > 0) a = cmul(a, tw[r * k * twm]);
                v[gi][r] = a;
            }
            dft<R>(v[gi]);
            dsti[gi] = (q * Ns * R + k) * T + t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
        if (dsti[gi] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) buf[dsti[gi] + r * Ns * T] = v[gi][r];
        }
    }
    __syncthreads();
}

// Full forward transform of T interleaved length-N transforms in place in
// buf (T * N float2), 256 threads; requires N * T <= 3200.  Ends with a
// barrier.
template <int T>
__device__ void fft_lds_inplace(float2 *buf, const FftPlan &plan, const float2 *tw)
{
    for (int p = 0; p < plan.npass; p++) {
        const int R = plan.radix[p], Ns = plan.ns[p];
        const unsigned mag = plan.mag[p];
        const int twm = plan.twm[p];
        switch (R) {
        case 2: stockham_pass_inplace<2, T, 7>(buf, plan.N, Ns, mag, twm, tw); break;
        case 3: stockham_pass_inplace<3, T, 5>(buf, plan.N, Ns, mag, twm, tw); break;
        case 4: stockham_pass_inplace<4, T, 4>(buf, plan.N, Ns, mag, twm, tw); break;
        case 5: stockham_pass_inplace<5, T, 3>(buf, plan.N, Ns, mag, twm, tw); break;
        case 6: stockham_pass_inplace<6, T, 3>(buf, plan.N, Ns, mag, twm, tw); break;
        case 7: stockham_pass_inplace<7, T, 2>(buf, plan.N, Ns, mag, twm, tw); break;
        case 11: stockham_pass_inplace<11, T, 2>(buf, plan.N, Ns, mag, twm, tw); break;
        case 13: stockham_pass_inplace<13, T, 1>(buf, plan.N, Ns, mag, twm, tw); break;
        default: stockham_pass_inplace<8, T, 2>(buf, plan.N, Ns, mag, twm, tw); break;
        }
    }
}

// Full forward transform (ping-pong between a and b); returns the buffer that
// holds the result.  Ends with a barrier.
template <int T>
__device__ float2 *fft_lds(float2 *a, float2 *b, const FftPlan &plan, const float2 *tw)
{
    for (int p = 0; p < plan.npass; p++) {
        const int R = plan.radix[p], Ns = plan.ns[p];
        const unsigned mag = plan.mag[p];
        const int twm = plan.twm[p];
        switch (R) {
        case 2: stockham_pass<2, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 3: stockham_pass<3, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 4: stockham_pass<4, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 5: stockham_pass<5, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 6: stockham_pass<6, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 7: stockham_pass<7, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 11: stockham_pass<11, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        case 13: stockham_pass<13, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        default: stockham_pass<8, T>(a, b, plan.N, Ns, mag, twm, tw); break;
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Compile-time plans for the common pads: the radices make_plan picks (8, 6,
// 4, 2, 3, 5, greedily), so a fixed-size transform computes exactly what the
// runtime plan computes, with the strides, divisions and loop bounds folded
// into constants.
// ---------------------------------------------------------------------------
constexpr int plan_radix(int N, int p)
{
    const int order[9] = {8, 6, 4, 2, 3, 5, 7, 11, 13};
    int m = N, k = 0;
    for (int oi = 0; oi < 9; oi++)
        while (m % order[oi] == 0) {
            if (k == p) return order[oi];
            k++;
            m /= order[oi];
        }
    return 0;
}
constexpr int plan_ns(int N, int p)
{
    int ns = 1;
    for (int i = 0; i < p; i++) ns *= plan_radix(N, i);
    return ns;
}

// In-place pass of a fixed plan (256 threads): reads and transforms every
// butterfly of the thread into registers, one barrier, writes them back.
template <int R, int T, int N, int NS>
__device__ __forceinline__ void stockham_pass_fixed(float2 *buf, const float2 *__restrict__ tw)
{
    constexpr int NR = N / R, TOTAL = NR * T, G = (TOTAL + 255) / 256, TWM = N / (NS * R);
    float2 v[G][R];
    int dsti[G];
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
        const int b = threadIdx.x + gi * 256;
        dsti[gi] = -1;
        if (TOTAL % 256 == 0 || b < TOTAL) {
            const int t = b % T;
            const int j = b / T;
            const int q = j / NS;
            const int k = j - q * NS;
#pragma unroll
            for (int r = 0; r < R; r++) {
                float2 a = buf[(j + r * NR) * T + t];
                if (NS > 1 && r > 0) a = cmul(a, tw[r * k * TWM]);  // the first pass has k = 0: unit twiddles
                v[gi][r] = a;
            }
            dft<R>(v[gi]);
            dsti[gi] = (q * NS * R + k) * T + t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
        if (dsti[gi] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) buf[dsti[gi] + r * NS * T] = v[gi][r];
        }
    }
    __syncthreads();
}

// Full forward transform of T interleaved length-N transforms in place
// (fixed plan, 256 threads).  Ends with a barrier.
template <int T, int N, int P = 0>
__device__ __forceinline__ void fft_fixed(float2 *buf, const float2 *tw)
{
    if constexpr (plan_radix(N, P) != 0) {
        stockham_pass_fixed<plan_radix(N, P), T, N, plan_ns(N, P)>(buf, tw);
        fft_fixed<T, N, P + 1>(buf, tw);
    }
}

// In place when the tile fits the register bound of fft_lds_inplace, else
// ping-pong through b (the launch sizes the LDS accordingly).
template <int T>
__device__ __forceinline__ float2 *fft_tile(float2 *a, float2 *b, const FftPlan &plan, const float2 *tw)
{
    if (plan.N * T <= 3200) {
        fft_lds_inplace<T>(a, plan, tw);
        return a;
    }
    return fft_lds<T>(a, b, plan, tw);
}

__device__ __forceinline__ void load_twiddles(float2 *tw, const float2 *__restrict__ g, int N)
{
    for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = g[i];
}

}  // namespace fh

// Fourier-space gridding kernels of the "fourier" reconstructor path.
//
// Per-pixel semantics restate FourierReconstructor::do_insert_slice_work
// (libEM/reconstructor.cpp:1122-1210), FourierInserter3DMode2 / Mode2l
// ::insert_pixel (libEM/reconstructor_tools.cpp:128-239), Util::fast_exp /
// hypot3sq (libEM/util.cpp:788-800, util.h:816-819) and
// EMData::add_complex_at_fast (libEM/emdata_core.h:368-413).  The float
// arithmetic of coordinates and weights is evaluated in the reference's
// order with contraction off, so every contribution is bit-identical to the
// reference's; only the order of the FP32 sums differs.
//
//  k_insert_plain    one launch per (slice, symmetry op), one thread per pixel,
//                    global atomics per contribution (the plain form).
//  k_bin_*           bins (slice, brick) pairs over the unfolded frequency
//                    grid [-P/2, P/2]^3 cut into 8^3 bricks, keyed by brick
//                    and the slice's dominant normal axis.
//  k_scatter_insert  one wave per work item holds the brick's accumulator in
//                    LDS and adds the stencils of the slice pixels reaching
//                    the brick with plain read-modify-writes (lanes sharing
//                    a stencil origin take turns, elected by an LDS tag).
//                    Each item writes its partial brick to HBM.
//  k_merge_partials  one thread per half-volume voxel sums the partial
//                    bricks of its unfolded source cells under the Hermitian
//                    fold of add_complex_at_fast and adds them with plain
//                    stores (each voxel has one owner: no atomics).
#pragma clang fp contract(off)
#include "fh_device.hpp"

namespace fh {

#define EXPN 400  // fast_exp entries reachable by a 2x2x2 stencil (r^2 <= 3)

__device__ __forceinline__ void gadd(float *p, float v) { unsafeAtomicAdd(p, v); }

// ------------------------------------------------------------ compose
// t3d[q] = R[p] * S[s] with q = p*nsym + s (Transform operator*,
// libEM/transform.cpp:1384-1395, FP32 in source order).
__global__ void k_compose(const float *__restrict__ rots, const float *__restrict__ weights, int n,
                          const float *__restrict__ syms, int nsym, float *__restrict__ t3d,
                          float *__restrict__ wq)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n * nsym) return;
    const int p = q / nsym, s = q - p * nsym;
    const float *a = rots + 9 * (size_t)p;
    const float *b = syms + 9 * s;
    float *r = t3d + 9 * (size_t)q;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r[3 * i + j] = a[3 * i + 0] * b[0 * 3 + j] + a[3 * i + 1] * b[1 * 3 + j] + a[3 * i + 2] * b[2 * 3 + j];
    wq[q] = weights[p];
}

// The 2-D part of preprocess_slice (libEM/reconstructor.cpp:1031-1035): the
// "xform" processor's real-space resampling (TransformProcessor::transform,
// libEM/processor.cpp:11025-11075) of the centre zero-padded, edge-mean
// subtracted particle.  Output pixel (i, j) samples the padded image at
// inv . (i - P/2, j - P/2) + (P/2, P/2) bilinearly (Util::bilinear_interpolate,
// libEM/util.h:543-547; the last row/column reuse their own pixel), 0 outside.
// xf holds inv's m00 m01 m03 m10 m11 m13 per particle; float operation
// order as the reference (no contraction in this unit).
__global__ void __launch_bounds__(256) k_xform2d(const float *__restrict__ imgs, int B, int P, int off,
                                                 const float *__restrict__ means, const float *__restrict__ xf,
                                                 float *__restrict__ out)
{
    const int p = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * P) return;
    const int i = idx % P, j = idx / P;
    const float *m = xf + 6 * (size_t)p;
    const float x = (float)(i - P / 2), y = (float)(j - P / 2);
    const float x2 = (m[0] * x + m[1] * y + m[2]) + (float)(P / 2);
    const float y2 = (m[3] * x + m[4] * y + m[5]) + (float)(P / 2);
    float v = 0.0f;
    if (!(x2 < 0 || x2 >= (float)P || y2 < 0 || y2 >= (float)P)) {
        const int ii = (int)x2, jj = (int)y2;  // fast_floor of a non-negative value
        const int i1 = ii == P - 1 ? ii : ii + 1, j1 = jj == P - 1 ? jj : jj + 1;
        const float *d = imgs + (size_t)p * B * B;
        const float mean = means[p];
        // the padded image: particle pixel minus its edge mean inside the box, 0 outside
        auto px = [&](int a, int b) -> float {
            const int xa = a - off, yb = b - off;
            return ((unsigned)xa < (unsigned)B && (unsigned)yb < (unsigned)B) ? d[(size_t)yb * B + xa] - mean : 0.0f;
        };
        const float p1 = px(ii, jj), p2 = px(i1, jj), p3 = px(ii, j1), p4 = px(i1, j1);
        const float t = x2 - (float)ii, u = y2 - (float)jj;
        v = (1 - t) * (1 - u) * p1 + t * (1 - u) * p2 + (1 - t) * u * p3 + t * u * p4;
    }
    out[(size_t)p * P * P + idx] = v;
}

hipError_t launch_xform2d(const float *imgs, int n, int B, int P, const float *means, const float *xf, float *out,
                          hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int d = B - P;
    const int fl = d >= 0 ? d / 2 : -((-d + 1) / 2);  // clip_inplace Region((B-P)//2, ...)
    hipLaunchKernelGGL(k_xform2d, dim3((P * P + 255) / 256, n), dim3(256), 0, s, imgs, B, P, -fl, means, xf, out);
    return hipGetLastError();
}

hipError_t launch_compose(const float *rots, const float *weights, int n, const float *syms, int nsym,
                          float *t3d, float *wq, hipStream_t s)
{
    const int tot = n * nsym;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compose, dim3((tot + 255) / 256), dim3(256), 0, s, rots, weights, n, syms, nsym, t3d, wq);
    return hipGetLastError();
}

// ------------------------------------------------------------ per-pixel math
// (rx, ry, 0) * t3d, then x (nx-2, ny, nz) = P (libEM/reconstructor.cpp:1171-1196)
__device__ __forceinline__ void pixel_coords(int x, int y, const float *m, const float *rtab, int H, float Pf,
                                             float &xx, float &yy, float &zz)
{
    const float rx = rtab[x + H], ry = rtab[y + H];
    xx = (rx * m[0] + ry * m[3]) * Pf;
    yy = (rx * m[1] + ry * m[4]) * Pf;
    zz = (rx * m[2] + ry * m[5]) * Pf;
}

// slice.get_complex_at(x, y) for x >= 0 (libEM/emdata_core.cpp:126-132);
// the x=0, y<0 pixel is read as the conjugate of (0, -y).
__device__ __forceinline__ float2 slice_value(const float2 *__restrict__ S, int x, int y, int P, int H)
{
    const bool cj = x == 0 && y < 0;
    const int row = cj ? -y : (y < 0 ? P + y : y);
    float2 v = S[row * (H + 1) + x];  // P (P/2 + 1) < 2^31
    if (cj) v.y = -v.y;
    return v;
}

// gauss_2 / trilinear weight of corner (i,j,k) for a pixel at (xx,yy,zz).
__device__ __forceinline__ float corner_weight(int mode, int i, int j, int k, float xx, float yy, float zz, float h,
                                               float weight, const float *etab)
{
    const float dx = (float)i - xx, dy = (float)j - yy, dz = (float)k - zz;
    if (mode == 0) {
        const float r = dx * dx + dy * dy + dz * dz;  // Util::hypot3sq
        const float f = r * h;                          // -(-r*h)
        const int g = (int)__builtin_fma((double)f, 50.0, 0.5);  // Util::fast_exp index: f*50 is exact in double
        return etab[g] * weight;
    }
    return (float)((1.0 - (double)fabsf(dx)) * (1.0 - (double)fabsf(dy)) * (1.0 - (double)fabsf(dz)) *
                   (double)weight);
}

__device__ __forceinline__ size_t cidx(int xs, int ys, int zs, int P, int H)
{
    return ((size_t)zs * P + ys) * (size_t)(H + 1) + xs;
}

// EMData::add_complex_at_fast with the weight update norm[off/2] += gg of
// insert_pixel, as device atomics.  (x,y,z) unfolded, value v, weight gg.
__device__ __forceinline__ void add_complex_atomic(float2 *data, float *norm, int P, int H, int x, int y, int z,
                                                   float vr, float vi, float gg)
{
    if (abs(x) >= H + 1 || abs(y) > H || abs(z) > H) return;
    float *d = reinterpret_cast<float *>(data);
    const int wy = y < 0 ? P + y : y, wz = z < 0 ? P + z : z;
    const int ny = y <= 0 ? -y : P - y, nz = z <= 0 ? -z : P - z;
    if (x == 0) {
        if (y == 0 && z == 0) {
            gadd(&d[0], vr);
            d[1] = 0.f;
            gadd(&norm[0], gg);
            return;
        }
        const size_t i2 = cidx(0, ny, nz, P, H);
        gadd(&d[2 * i2], vr);
        gadd(&d[2 * i2 + 1], -vi);
    }
    if (abs(x) == H) {
        if (y == 0 && z == 0) {
            gadd(&d[2 * H], vr);
            d[2 * H + 1] = 0.f;
            gadd(&norm[H], gg);
            return;
        }
        const size_t i2 = cidx(H, ny, nz, P, H);
        gadd(&d[2 * i2], vr);
        gadd(&d[2 * i2 + 1], -vi);
    }
    if (x < 0) {
        const size_t i = cidx(-x, ny, nz, P, H);
        gadd(&d[2 * i], vr);
        gadd(&d[2 * i + 1], -vi);
        gadd(&norm[i], gg);
        return;
    }
    const size_t i = cidx(x, wy, wz, P, H);
    gadd(&d[2 * i], vr);
    gadd(&d[2 * i + 1], vi);
    gadd(&norm[i], gg);
}

// ------------------------------------------------------------ other inserter modes
// Util::fast_exp (util.cpp:788-800) with its full 1000-entry table
__device__ __forceinline__ float fast_exp_full(float f, const float *__restrict__ etab)
{
    return (f > 0 || f < -19.98) ? expf(f) : etab[(int)(-(double)f * 50.0 + 0.5)];
}

// e^-|x| I0(x), the exponentially scaled modified Bessel function of order 0
// that kaiser_bessel takes from GSL (gsl_sf_bessel_i0_scaled): the
// polynomial approximations of Abramowitz & Stegun 9.8.1 (|x| <= 3.75,
// |error| < 1.6e-7 of I0) and 9.8.2 (x > 3.75, |error| < 1.9e-7 of
// sqrt(x) e^-x I0(x)), evaluated in double.
__host__ __device__ inline double bessel_i0_scaled(double x)
{
    const double ax = fabs(x);
    if (ax <= 3.75) {
        const double t = (ax / 3.75) * (ax / 3.75);
        const double i0 = 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 +
                          t * (0.0360768 + t * 0.0045813)))));
        return i0 * exp(-ax);
    }
    const double u = 3.75 / ax;
    return (0.39894228 + u * (0.01328592 + u * (0.00225319 + u * (-0.00157565 + u * (0.00916281 +
            u * (-0.02057706 + u * (0.02635537 + u * (-0.01647633 + u * 0.00392377)))))))) / sqrt(ax);
}

// The inserters the brick scatter does not take, in the global-atomic form
// (libEM/reconstructor_tools.cpp):
//   mode 2 gauss_var (Mode6, :382-433): 6x6x6 stencil from floor(p - 2.5),
//     h = 1 / (|p|^2 / 4000 + 0.15) at least 0.1, weight fast_exp(-r h) w;
//   mode 3 gauss_5 (Mode5): the same stencil with h = 1 / I5G;
//   mode 4 nearest_neighbor (Mode1, :113-126): the voxel floor(p + 0.5);
//   mode 5 gauss_3 (Mode3): 3x3x3 from floor(p - 0.5), h = 32 / ((8 + |p|) I3G),
//     weight fast_exp(-r h) w / (1 + 6 e^-h + 12 e^-2h + 8 e^-3h);
//   mode 10 fourier_iter (FourierIterReconstructor::insert_slice,
//     libEM/reconstructor.cpp:520-637): gauss_var's stencil and width with x
//     clamped to |x| <= P/2 - 1 (nx2 = nx/2 - 2); mode 11 the same
//     reconstructor with a reference volume: value dt + ref(i,j,k) -
//     trilinear ref(p), h = 1/2, cells inside i^2 + j^2 + k^2 <= P^2/4.
// Stencils are clamped to [-P/2, P/2] per axis; subnormal weights are
// dropped (flush-to-zero: the reference's -ffast-math build and the float
// atomics do not keep them).
__device__ void insert_wide(int mode, float2 dt, float xx, float yy, float zz, float w, int P, int H,
                            const float *__restrict__ etab, float2 *data, float *norm, const float2 *ref);

// ------------------------------------------------------------ plain form
// Util::hypot_fast_int(x, y) (libEM/util.cpp:764-785): round(hypot(x, y)),
// exact for 0 <= x^2 + y^2 < 2^24 (sqrt of an integer is never a half-integer).
__device__ __forceinline__ int hypot_int(int x, int y)
{
    const int n = x * x + y * y;
    int r = (int)__builtin_amdgcn_sqrtf((float)n);
    r -= (r * r > n);
    r += ((r + 1) * (r + 1) <= n);
    return r + (n > r * r + r);
}

__global__ void __launch_bounds__(256) k_insert_plain(const float2 *__restrict__ S, InsertConst c, Mat9 m,
                                                      float weight, float2 *data, float *norm,
                                                      const float *__restrict__ exptab,
                                                      const float *__restrict__ rtab_g,
                                                      const float *__restrict__ rw)
{
    const int H = c.H, P = c.P;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = (int)blockIdx.y - H;
    if (x > H) return;
    if (!c.corners && x * x + y * y > c.rcut * c.rcut + c.rcut) return;  // hypot_fast_int(x,y) > rcut
    // class_ssnr weighting (reconstructor.cpp:1170-1184): the weight of radius r
    if (rw) weight = rw[hypot_int(x, y)];
    float xx, yy, zz;
    pixel_coords(x, y, m.m, rtab_g, H, c.Pf, xx, yy, zz);
    if (c.mode >= 2) {
        insert_wide(c.mode, slice_value(S, x, y, P, H), xx, yy, zz, weight, P, H, exptab, data, norm, c.ref);
        return;
    }
    const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
    if (x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H) return;
    const float2 dt = slice_value(S, x, y, P, H);
    for (int k = z0; k <= z0 + 1; k++)
        for (int j = y0; j <= y0 + 1; j++)
            for (int i = x0; i <= x0 + 1; i++) {
                const float gg = corner_weight(c.mode, i, j, k, xx, yy, zz, c.h, weight, exptab);
                add_complex_atomic(data, norm, P, H, i, j, k, dt.x * gg, dt.y * gg, gg);
            }
}

// The plain (global-atomic) form over every slice of a batch in one launch:
// block (xb, q) covers pixels x in [64 xb, 64 xb + 64) of rows y of slice q
// = p nsym + s, with t3d[q] and weight wq[q] (or the particle's class_ssnr
// radial weights).  Used for the inserter modes the brick scatter does not
// take (gauss_var's 6x6x6 stencil).
__global__ void __launch_bounds__(64) k_insert_plain_batch(const float2 *__restrict__ slices, int nsym, int nxb,
                                                           const float *__restrict__ t3d, const float *__restrict__ wq,
                                                           InsertConst c, float2 *data, float *norm,
                                                           const float *__restrict__ exptab,
                                                           const float *__restrict__ rtab_g,
                                                           const float *__restrict__ rw)
{
    const int H = c.H, P = c.P;
    const int q = blockIdx.x / nxb;
    const int x = (blockIdx.x - q * nxb) * 64 + threadIdx.x;
    const int y = (int)blockIdx.y - H;
    const float wgt = wq[q];
    if (x > H || wgt == 0.0f) return;
    if (!c.corners && x * x + y * y > c.rcut * c.rcut + c.rcut) return;  // hypot_fast_int(x,y) > rcut
    const int p = q / nsym;
    const float weight = rw ? rw[(size_t)p * (H + 1) + hypot_int(x, y)] : wgt;
    const float2 *S = slices + (size_t)p * P * (H + 1);
    const float *m = t3d + 9 * (size_t)q;
    float xx, yy, zz;
    pixel_coords(x, y, m, rtab_g, H, c.Pf, xx, yy, zz);
    const float2 dt = slice_value(S, x, y, P, H);
    if (c.mode >= 2) {
        insert_wide(c.mode, dt, xx, yy, zz, weight, P, H, exptab, data, norm, c.ref);
        return;
    }
    const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
    if (x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H) return;
    for (int k = z0; k <= z0 + 1; k++)
        for (int j = y0; j <= y0 + 1; j++)
            for (int i = x0; i <= x0 + 1; i++) {
                const float gg = corner_weight(c.mode, i, j, k, xx, yy, zz, c.h, weight, exptab);
                add_complex_atomic(data, norm, P, H, i, j, k, dt.x * gg, dt.y * gg, gg);
            }
}

// The wide modes with LDS accumulation: a workgroup takes an 8 x 8 tile of
// one slice's pixels (4 threads per pixel, 54 of the 216 stencil cells each),
// adds every contribution into an LDS window of 17^3 unfolded cells with LDS
// float atomics, then adds each touched cell to the volume once through the
// add_complex_at_fast fold (global atomics): about ten times fewer global
// atomics than one per contribution.  Each contribution's value and weight
// are computed exactly as in insert_wide.  Across an 8 x 8 tile a pixel
// coordinate moves by at most 7 sqrt 2 (orthonormal slice axes), so the
// stencil origins of a tile span at most 10 cells per axis and the stencils
// 10 + (n - 1) cells for an n-cell stencil: within the 17-cell window for
// n <= 7 (gauss_var, gauss_5, gauss_3, kaiser_bessel, gridding_5,
// gridding_7); kaiser_bessel_derived's 9-cell stencil can reach 18, and the
// cells of a tile beyond the window go straight to the volume with global
// atomics (same result, less of the window's saving).  The window (59 KB)
// and the pixel records stay within 64 KB of LDS per workgroup.
constexpr int WT_TILE = 8;
constexpr int WT_WIN = 17;
constexpr int WT_CELLS = WT_WIN * WT_WIN * WT_WIN;

// gridding_5 / gridding_7 kernels, 3x oversampled (reconstructor_tools.cpp:455, :817)
#include "gridding_tables.inc"
__constant__ float c_grid5[9 * 9 * 9] = {FH_GRIDDING5_VALUES};
__constant__ float c_grid7[12 * 12 * 12] = {FH_GRIDDING7_VALUES};

// Util::fast_floor (libEM/util.h:874-880): one less than floor at negative integers
__device__ __forceinline__ int ref_fast_floor(float x) { return x < 0 ? (int)x - 1 : (int)x; }

// table index of one axis: |fast_floor((i - p) * 3 + 0.5)|, the 0.5 added in double as in the reference
__device__ __forceinline__ int grid_index(int i, float p, int n)
{
    const int t = abs(ref_fast_floor((float)((double)(((float)i - p) * 3.0f) + 0.5)));
    return min(t, n - 1);  // <= n - 1 for every finite p (the stencil spans |i - p| <= n / 3 - 0.5)
}

struct WidePix {
    float xx, yy, zz, dre, dim, w, h, ww;
    float nre, nim;        // fourier_iter with a reference: the reference trilinear at p
    int lo[3], hi[3];      // stencil cells, clamped to the volume
    int a[3];              // stencil origin before the clamp
    int H;
    const float2 *ref;
};

// EMData::get_complex_at(x, y, z) of a reference volume in the reconstructor's
// layout (libEM/emdata_core.cpp:135-144): 0 outside, conjugate for x < 0
__device__ __forceinline__ float2 ref_get(const float2 *__restrict__ ref, int x, int y, int z, int H)
{
    const int P = 2 * H;
    if (abs(x) >= H + 1 || abs(y) > H || abs(z) > H) return make_float2(0.f, 0.f);
    if (x < 0) {
        const float2 v = ref[cidx(-x, y <= 0 ? -y : P - y, z <= 0 ? -z : P - z, P, H)];
        return make_float2(v.x, -v.y);
    }
    return ref[cidx(x, y < 0 ? P + y : y, z < 0 ? P + z : z, P, H)];
}

// stencil range (clamped to [-P/2, P/2]) and width of one pixel; false if the
// reference's range check rejects it (insert_wide's first half)
__device__ __forceinline__ bool wide_setup(int mode, float xx, float yy, float zz, float w, int H,
                                           const float *__restrict__ etab, WidePix &px,
                                           const float2 *ref = nullptr)
{
    px.xx = xx;
    px.yy = yy;
    px.zz = zz;
    px.w = w;
    px.ww = w;
    px.h = 0.f;
    px.H = H;
    px.ref = ref;
    px.nre = px.nim = 0.f;
    if (mode >= 10) {
        // FourierIterReconstructor::insert_slice: floor(p - 2.5) + [0, 5]; rejected when the
        // origin is below -n2 - 4 or above n2 + 3, n2 = P/2 - 1 for x and P/2 for y, z; clamped to [-n2, n2]
        const float pp[3] = {xx, yy, zz};
        for (int d = 0; d < 3; d++) {
            const int n2 = d == 0 ? H - 1 : H;
            const int a = (int)floor((double)pp[d] - 2.5);
            if (a < -n2 - 4 || a > n2 + 3) return false;
            px.a[d] = a;
            px.lo[d] = max(a, -n2);
            px.hi[d] = min(a + 5, n2);
            if (px.hi[d] < px.lo[d]) return false;
        }
        if (mode == 10) {
            float h = (float)(1.0f / ((double)((xx * xx + yy * yy + zz * zz) / 4000) + .15));
            px.h = h < 0.1 ? (float)0.1 : h;
            return true;
        }
        // Util::trilinear_interpolate_complex (libEM/util.h:645-654) at floor(p)
        px.h = 1.0f / 2.0f;
        const float xx0 = floorf(xx), yy0 = floorf(yy), zz0 = floorf(zz);
        const int x0 = (int)xx0, y0 = (int)yy0, z0 = (int)zz0;
        const float t = xx - xx0, u = yy - yy0, v = zz - zz0;
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const float2 q = ref_get(ref, x0 + (c & 1), y0 + ((c >> 1) & 1), z0 + (c >> 2), H);
            const float a = (c & 1) ? t : 1 - t, b = (c & 2) ? u : 1 - u, e = (c & 4) ? v : 1 - v;
            const float wc = a * b * e;
            if (c == 0) {
                ar = wc * q.x;
                ai = wc * q.y;
            } else {
                ar = ar + wc * q.x;
                ai = ai + wc * q.y;
            }
        }
        px.nre = ar;
        px.nim = ai;
        return true;
    }
    if (mode == 4) {
        const int c0 = (int)floor((double)(xx + 0.5f)), c1 = (int)floor((double)(yy + 0.5f));
        const int c2 = (int)floor((double)(zz + 0.5f));
        px.lo[0] = px.hi[0] = px.a[0] = c0;
        px.lo[1] = px.hi[1] = px.a[1] = c1;
        px.lo[2] = px.hi[2] = px.a[2] = c2;
        return true;
    }
    if (mode == 9) {
        // kaiser_bessel_derived: floor(p - 4) + [0, 8]; h = ws[0..2] = I0s(0) / I0s(32 pi) (ws[3..4] = 1),
        // ww = their float sum in the reference's order
        const float pp[3] = {xx, yy, zz};
        for (int d = 0; d < 3; d++) {
            const int a = (int)floorf(pp[d] - 4.0f);
            if (a < -H - 7 || a > H + 6) return false;
            px.a[d] = a;
            px.lo[d] = max(a, -H);
            px.hi[d] = min(a + 8, H);
            if (px.hi[d] < px.lo[d]) return false;
        }
        const float w0 = (float)(bessel_i0_scaled(0.0) / bessel_i0_scaled(M_PI * 32.0));
        px.h = w0;
        px.ww = w0 + w0 + w0 + 1.0f + 1.0f;
        return true;
    }
    if (mode == 7 || mode == 8) {
        // gridding_5: fast_floor(p - 1.5) + [0, 4]; gridding_7: fast_floor(p - 2.5) + [0, 6]
        const double sh = mode == 8 ? 2.5 : 1.5;
        const int span = mode == 8 ? 6 : 4, lo = mode == 8 ? 6 : 4, hi = mode == 8 ? 5 : 3;
        const float pp[3] = {xx, yy, zz};
        for (int d = 0; d < 3; d++) {
            const int a = ref_fast_floor((float)((double)pp[d] - sh));
            if (a < -H - lo || a > H + hi) return false;
            px.a[d] = a;
            px.lo[d] = max(a, -H);
            px.hi[d] = min(a + span, H);
            if (px.hi[d] < px.lo[d]) return false;
        }
        return true;
    }
    // gauss_3: floor(p - .5) + [0, 2]; kaiser_bessel: floor(p - 2) + [0, 4]; else floor(p - 2.5) + [0, 5]
    const double off = mode == 5 ? 0.5 : 2.5;
    const int span = mode == 5 ? 2 : mode == 6 ? 4 : 5, lo = mode == 5 ? 2 : mode == 6 ? 7 : 4;
    const int hi = mode == 5 ? 1 : mode == 6 ? 6 : 3;
    const float pp[3] = {xx, yy, zz};
    for (int d = 0; d < 3; d++) {
        const int a = mode == 6 ? (int)floor((double)(pp[d] - 2.0f)) : (int)floor((double)pp[d] - off);
        if (a < -H - lo || a > H + hi) return false;
        px.a[d] = a;
        px.lo[d] = max(a, -H);
        px.hi[d] = min(a + span, H);
        if (px.hi[d] < px.lo[d]) return false;  // an empty loop in the reference
    }
    if (mode == 6) {
        px.h = (float)(H - 1);  // nx2 - 1 of the kaiser_bessel radius scaling
    } else if (mode == 2) {
        float h = (float)(1.0f / ((double)((xx * xx + yy * yy + zz * zz) / 4000) + .15));
        px.h = h < 0.1 ? (float)0.1 : h;
    } else if (mode == 3) {
        px.h = 1.0f / (float)(10.4 / (M_PI * M_PI));
    } else {
        const float I3G = (float)(6.4 / (M_PI * M_PI));
        px.h = 32.0f / ((8.0f + sqrt_rn(xx * xx + yy * yy + zz * zz)) * I3G);
        px.ww = w / (1.0f + 6.0f * fast_exp_full(-px.h, etab) + 12 * fast_exp_full(-px.h * 2.0f, etab) +
                     8 * fast_exp_full(-px.h * 3.0f, etab));
    }
    return true;
}

// value and weight of cell (i, j, k) for the pixel (insert_wide's loop body);
// false when the contribution is dropped
__device__ __forceinline__ bool wide_term(int mode, const WidePix &px, int i, int j, int k,
                                          const float *__restrict__ etab, float &vr, float &vi, float &g)
{
    if (mode == 4) {
        vr = px.dre * px.w;
        vi = px.dim * px.w;
        g = px.w;
        return true;
    }
    if (mode == 11) {
        // updt = dt + ref(i,j,k) - ref~(p) under fast_exp(-r / 2), inside the sphere of radius P/2
        if ((long long)i * i + (long long)j * j + (long long)k * k > (long long)(2 * px.H) * (2 * px.H) / 4)
            return false;
        const float2 nt = ref_get(px.ref, i, j, k, px.H);
        const float ur = (px.dre + nt.x) - px.nre, ui = (px.dim + nt.y) - px.nim;
        const float dx = (float)i - px.xx, dy = (float)j - px.yy, dz = (float)k - px.zz;
        const float gg = fast_exp_full(-(dx * dx + dy * dy + dz * dz) * px.h, etab);
        if (fabsf(gg * px.w) < FLT_MIN) return false;
        vr = ur * gg * px.w;
        vi = ui * gg * px.w;
        g = gg * px.w;
        return true;
    }
    if (mode == 9) {
        // Mode10 (reconstructor_tools.cpp:759-812): kb = ws[0] + ... + ws[q], q <= min(r, 4) with r the
        // squared distance; value dt sqrt(kb / sum ws) w, weight w
        const float dx = (float)i - px.xx, dy = (float)j - px.yy, dz = (float)k - px.zz;
        const float r = dx * dx + dy * dy + dz * dz;
        const float lim = r < 4.0f ? r : 4.0f;
        float kb = 0.0f;
        for (int q = 0; q <= 4 && (float)q <= lim; q++) kb += q < 3 ? px.h : 1.0f;
        const float dn = sqrt_rn(__fdiv_rn(kb, px.ww));
        vr = px.dre * dn * px.w;
        vi = px.dim * dn * px.w;
        g = px.w;
        return true;
    }
    if (mode == 7 || mode == 8) {
        // Mode7 / Mode11: value dt gg w, weight w (reconstructor_tools.cpp:539-576, 964-1001)
        const int n = mode == 8 ? 12 : 9;
        const float *tab = mode == 8 ? c_grid7 : c_grid5;
        const float gg = tab[(grid_index(i, px.xx, n) * n + grid_index(j, px.yy, n)) * n + grid_index(k, px.zz, n)];
        vr = px.dre * gg * px.w;
        vi = px.dim * gg * px.w;
        g = px.w;
        return true;
    }
    const float dx = (float)i - px.xx, dy = (float)j - px.yy, dz = (float)k - px.zz;
    if (mode == 6) {
        // Mode9 (reconstructor_tools.cpp): kb = I0s(pi a sqrt(1 - (r / (nx2 - 1) - 1)^2)) / I0s(pi a),
        // a = 15, r the squared distance; value dt * kb * w, weight w
        const float r = dx * dx + dy * dy + dz * dz;
        const float t = r / px.h - 1;
        const float a = 15.0f;
        const float kb = (float)(bessel_i0_scaled(M_PI * a * (double)sqrt_rn(1.0f - t * t)) /
                                 bessel_i0_scaled(M_PI * a));
        vr = px.dre * kb * px.w;
        vi = px.dim * kb * px.w;
        g = px.w;
        return true;
    }
    const float gg = fast_exp_full(-(dx * dx + dy * dy + dz * dz) * px.h, etab);
    if (mode == 5) {
        g = gg * px.ww;
        if (fabsf(g) < FLT_MIN) return false;
        vr = px.dre * g;
        vi = px.dim * g;
        return true;
    }
    if (fabsf(gg * px.ww) < FLT_MIN) return false;
    vr = px.dre * gg * px.ww;
    vi = px.dim * gg * px.ww;
    g = gg * px.ww;
    return true;
}

// one pixel's stencil straight into the volume (the single-slice and
// fallback form of the wide modes), the same terms as the LDS-window form
__device__ void insert_wide(int mode, float2 dt, float xx, float yy, float zz, float w, int P, int H,
                            const float *__restrict__ etab, float2 *data, float *norm, const float2 *ref)
{
    WidePix px;
    if (!wide_setup(mode, xx, yy, zz, w, H, etab, px, ref)) return;
    px.dre = dt.x;
    px.dim = dt.y;
    for (int k = px.lo[2]; k <= px.hi[2]; k++)
        for (int j = px.lo[1]; j <= px.hi[1]; j++)
            for (int i = px.lo[0]; i <= px.hi[0]; i++) {
                float vr, vi, g;
                if (wide_term(mode, px, i, j, k, etab, vr, vi, g)) add_complex_atomic(data, norm, P, H, i, j, k, vr, vi, g);
            }
}

__global__ void __launch_bounds__(256) k_insert_wide_tiles(const float2 *__restrict__ slices, int nsym, int ntx,
                                                           int nty, const float *__restrict__ t3d,
                                                           const float *__restrict__ wq, InsertConst c, float2 *data,
                                                           float *norm, const float *__restrict__ exptab,
                                                           const float *__restrict__ rtab_g,
                                                           const float *__restrict__ rw)
{
    extern __shared__ float win[];  // re, im, weight: 3 x WT_CELLS
    __shared__ WidePix pix[WT_TILE * WT_TILE];
    __shared__ int okp[WT_TILE * WT_TILE];
    __shared__ int wlo[3], nvalid;
    const int H = c.H, P = c.P;
    const int tiles = ntx * nty;
    const int q = blockIdx.x / tiles;
    const int tt = blockIdx.x - q * tiles;
    const int tx0 = (tt % ntx) * WT_TILE, ty0 = (tt / ntx) * WT_TILE - H;
    const int p = q / nsym;
    const float wgt = wq[q];
    if (threadIdx.x < 3) wlo[threadIdx.x] = INT_MAX;
    if (threadIdx.x == 0) nvalid = 0;
    __syncthreads();
    if (threadIdx.x < WT_TILE * WT_TILE) {
        const int x = tx0 + (threadIdx.x % WT_TILE), y = ty0 + (threadIdx.x / WT_TILE);
        bool ok = wgt != 0.0f && x <= H && y <= H - 1 && (c.corners || x * x + y * y <= c.rcut * c.rcut + c.rcut);
        if (ok) {
            const float weight = rw ? rw[(size_t)p * (H + 1) + hypot_int(x, y)] : wgt;
            float xx, yy, zz;
            pixel_coords(x, y, t3d + 9 * (size_t)q, rtab_g, H, c.Pf, xx, yy, zz);
            WidePix px;
            ok = wide_setup(c.mode, xx, yy, zz, weight, H, exptab, px, c.ref);
            if (ok) {
                const float2 dt = slice_value(slices + (size_t)p * P * (H + 1), x, y, P, H);
                px.dre = dt.x;
                px.dim = dt.y;
                pix[threadIdx.x] = px;
                for (int d = 0; d < 3; d++) atomicMin(&wlo[d], px.lo[d]);
                atomicAdd(&nvalid, 1);
            }
        }
        okp[threadIdx.x] = ok;
    }
    __syncthreads();
    if (nvalid == 0) return;  // block-uniform
    for (int i = threadIdx.x; i < 3 * WT_CELLS; i += blockDim.x) win[i] = 0.f;
    __syncthreads();
    const int pi = threadIdx.x >> 2, part = threadIdx.x & 3;
    if (okp[pi]) {
        const WidePix px = pix[pi];
        const int nx = px.hi[0] - px.lo[0] + 1, ny = px.hi[1] - px.lo[1] + 1;
        const int ncell = nx * ny * (px.hi[2] - px.lo[2] + 1);
        for (int ci = part; ci < ncell; ci += 4) {
            const int kk = ci / (nx * ny), rem = ci - kk * nx * ny, jj = rem / nx, ii = rem - jj * nx;
            const int i = px.lo[0] + ii, j = px.lo[1] + jj, k = px.lo[2] + kk;
            float vr, vi, g;
            if (!wide_term(c.mode, px, i, j, k, exptab, vr, vi, g)) continue;
            const int a = i - wlo[0], b = j - wlo[1], e = k - wlo[2];
            if ((unsigned)a < (unsigned)WT_WIN && (unsigned)b < (unsigned)WT_WIN && (unsigned)e < (unsigned)WT_WIN) {
                const int cell = (e * WT_WIN + b) * WT_WIN + a;
                atomicAdd(&win[cell], vr);
                atomicAdd(&win[WT_CELLS + cell], vi);
                atomicAdd(&win[2 * WT_CELLS + cell], g);
            } else {
                add_complex_atomic(data, norm, P, H, i, j, k, vr, vi, g);
            }
        }
    }
    __syncthreads();
    for (int cell = threadIdx.x; cell < WT_CELLS; cell += blockDim.x) {
        const float g = win[2 * WT_CELLS + cell];
        if (g == 0.f) continue;
        const int a = cell % WT_WIN, b = (cell / WT_WIN) % WT_WIN, e = cell / (WT_WIN * WT_WIN);
        add_complex_atomic(data, norm, P, H, wlo[0] + a, wlo[1] + b, wlo[2] + e, win[cell], win[WT_CELLS + cell], g);
    }
}

hipError_t launch_insert_plain_batch(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                                     InsertConst c, float2 *data, float *norm, const float *exptab, const float *rtab,
                                     const float *rw, hipStream_t s)
{
    if (nslices <= 0) return hipSuccess;
    if (c.mode >= 2) {
        const int ntx = (c.H + 1 + WT_TILE - 1) / WT_TILE, nty = (c.P + WT_TILE - 1) / WT_TILE;
        const long long blocks = (long long)ntx * nty * nslices;
        if (blocks < INT32_MAX) {
            hipLaunchKernelGGL(k_insert_wide_tiles, dim3((unsigned)blocks), dim3(256), sizeof(float) * 3 * WT_CELLS, s,
                               slices, nsym, ntx, nty, t3d, wq, c, data, norm, exptab, rtab, rw);
            return hipGetLastError();
        }
    }
    const int nxb = (c.H + 1 + 63) / 64;
    hipLaunchKernelGGL(k_insert_plain_batch, dim3(nxb * nslices, c.P), dim3(64), 0, s, slices, nsym, nxb, t3d, wq, c,
                       data, norm, exptab, rtab, rw);
    return hipGetLastError();
}

hipError_t launch_insert_plain(const float2 *slice, InsertConst c, const float *t3d9, float weight, float2 *data,
                               float *norm, const float *exptab, const float *rtab, const float *rw, hipStream_t s)
{
    Mat9 m;
    for (int i = 0; i < 9; i++) m.m[i] = t3d9[i];
    dim3 grid((c.H + 1 + 63) / 64, c.P);
    hipLaunchKernelGGL(k_insert_plain, grid, dim3(64), 0, s, slice, c, m, weight, data, norm, exptab, rtab, rw);
    return hipGetLastError();
}

// ------------------------------------------------------------ slice agreement
// FourierReconstructor::do_compare_slice_work (libEM/reconstructor.cpp:1259-1381)
// with pixel_at (:1448-1461): nearest-voxel value/weight at each slice pixel
// x in [0, P/2+1] (the x = P/2+1 column reads the next row's x = 0 element,
// and past the last row 0), eight double sums per (slice, block).
__device__ __forceinline__ int round_ref(float v) { return v < 0 ? (int)(v - 0.5f) : (int)(v + 0.5f); }

__global__ void __launch_bounds__(256) k_agree(const float2 *__restrict__ slices, int nsym, const float *__restrict__ t3d,
                                               InsertConst c, float rmax, const float2 *__restrict__ data,
                                               const float *__restrict__ norm, const float *__restrict__ rtab,
                                               double *__restrict__ partial)
{
    const int H = c.H, P = c.P;
    const int q = blockIdx.y;
    const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = H + 2;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // dot vweight power power2 dotlow dlpow dlpow2 nval
    if (pidx < W * P) {
        const int x = pidx % W, y = pidx / W - H;
        const float *m = t3d + 9 * (size_t)q;
        if (!(x == 0 && y == 0)) {
            const float rx = rtab[x + H], ry = rtab[y + H];
            const float c0 = rx * m[0] + ry * m[3];
            const float c1 = rx * m[1] + ry * m[4];
            const float c2 = rx * m[2] + ry * m[5];
            if (!(fabsf(c0) > 0.5f || fabsf(c1) >= 0.5f || fabsf(c2) >= 0.5f)) {
                const float xx = c0 * c.Pf, yy = c1 * c.Pf, zz = c2 * c.Pf;
                const float *Sf = reinterpret_cast<const float *>(slices + (size_t)(q / nsym) * P * (H + 1));
                const int slen = (P + 2) * P;
                const int fidx = 2 * x + (P + 2) * (y < 0 ? P + y : y);
                const float d20 = fidx < slen ? Sf[fidx] : 0.f;
                const float d21 = fidx + 1 < slen ? Sf[fidx + 1] : 0.f;
                const int x0 = round_ref(xx), y0 = round_ref(yy), z0 = round_ref(zz);
                float vr = 0.f, vi = 0.f;
                size_t ci;
                if (x0 < 0) ci = cidx(-x0, y0 <= 0 ? -y0 : P - y0, z0 <= 0 ? -z0 : P - z0, P, H);
                else ci = cidx(x0, y0 < 0 ? P + y0 : y0, z0 < 0 ? P + z0 : z0, P, H);
                if (!(abs(x0) >= H + 1 || abs(y0) > H || abs(z0) > H)) {
                    const float2 v = data[ci];
                    vr = v.x;
                    vi = x0 < 0 ? -v.y : v.y;
                }
                const float nv = norm[ci];
                if (nv != 0.f) {
                    const float dt0 = __fdiv_rn(vr, nv), dt1 = __fdiv_rn(vi, nv), dt2 = nv;
                    const float r = sqrt_int_rn(x * x + y * y);  // Util::hypot_fast
                    const float dotv = (dt0 * d20 + dt1 * d21) * dt2;
                    const float pw = (dt0 * dt0 + dt1 * dt1) * dt2;
                    const float pw2 = (d20 * d20 + d21 * d21) * dt2;
                    if (r > 3 && r < rmax) {
                        acc[4] += dotv;
                        acc[5] += pw;
                        acc[6] += pw2;
                    }
                    acc[1] += dt2;
                    acc[0] += dotv;
                    acc[2] += pw;
                    acc[3] += pw2;
                    acc[7] += 1.0;
                }
            }
        }
    }
    __shared__ double red[8][256];
#pragma unroll
    for (int k = 0; k < 8; k++) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int k = 0; k < 8; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 8) partial[((size_t)q * gridDim.x + blockIdx.x) * 8 + threadIdx.x] = red[threadIdx.x][0];
}

int agree_blocks(int P) { return ((P / 2 + 2) * P + 255) / 256; }

// ---- batched agreement: compare against V - own(p) without mutating V ----
// own(p) at a storage voxel is what determine_slice_agreement's
// do_insert_slice_work(slice, R, -w) would subtract there
// (libEM/reconstructor.cpp:1236): the gauss_2 / trilinear contributions of
// all symmetry copies of particle p whose stencil cell folds onto the voxel.
// Unfolded sources of storage voxel (xs, ys, zs), following
// add_complex_at_fast: direct cells (xs, j, k) with wrap(j)=ys, wrap(k)=zs,
// conjugate cells (-xs, j, k) with wrapneg(j)=ys, wrapneg(k)=zs, and on the
// x=0 / x=P/2 planes the conjugate copies of the double insertion.

// contributions of slice copy m (t3d rows) of particle S to unfolded cell c
__device__ __forceinline__ void gather_cell(const float2 *__restrict__ S, const float *m, int ci, int cj, int ck,
                                            const InsertConst &c, float weight, const float *etab,
                                            const float *rtab, float &re, float &im, float &w)
{
    const int H = c.H, P = c.P;
    const float n0 = m[1] * m[5] - m[2] * m[4], n1 = m[2] * m[3] - m[0] * m[5], n2 = m[0] * m[4] - m[1] * m[3];
    const float dist = n0 * ci + n1 * cj + n2 * ck;
    if (fabsf(dist) > 1.7330f) return;  // |p - c|_inf < 1 on the plane => |n.c| < sqrt(3)
    const float s = m[0] * ci + m[1] * cj + m[2] * ck;
    const float t = m[3] * ci + m[4] * cj + m[5] * ck;
    const int xa = max(0, (int)ceilf(s - 1.7330f)), xb = min(H, (int)floorf(s + 1.7330f));
    const int ya = max(-H, (int)ceilf(t - 1.7330f)), yb = min(H - 1, (int)floorf(t + 1.7330f));
    for (int y = ya; y <= yb; y++)
        for (int x = xa; x <= xb; x++) {
            if (x * x + y * y > H * H + H) continue;  // radius cut (corners=false)
            float xx, yy, zz;
            pixel_coords(x, y, m, rtab, H, c.Pf, xx, yy, zz);
            const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
            if (x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H) continue;
            if ((unsigned)(ci - x0) > 1u || (unsigned)(cj - y0) > 1u || (unsigned)(ck - z0) > 1u) continue;
            const float gg = corner_weight(c.mode, ci, cj, ck, xx, yy, zz, c.h, weight, etab);
            const float2 dt = slice_value(S, x, y, P, H);
            re += dt.x * gg;
            im += dt.y * gg;
            w += gg;
        }
}

__device__ __forceinline__ void gather_cell_all(const float2 *S, const float *t3dp, int nsym, int ci, int cj, int ck,
                                                const InsertConst &c, float weight, const float *etab,
                                                const float *rtab, float &re, float &im, float &w)
{
    for (int s = 0; s < nsym; s++) gather_cell(S, t3dp + 9 * s, ci, cj, ck, c, weight, etab, rtab, re, im, w);
}

// own(p) at storage voxel (xs, ys, zs): value (ore, oim) and weight ow
__device__ void own_contribution(const float2 *S, const float *t3dp, int nsym, int xs, int ys, int zs,
                                 const InsertConst &c, float weight, const float *etab, const float *rtab,
                                 float &ore, float &oim, float &ow)
{
    const int H = c.H, P = c.P;
    ore = oim = ow = 0.f;
    int jd[2], njd = 0, jn[2], njn = 0, kd[2], nkd = 0, kn[2], nkn = 0;
    if (ys <= H) jd[njd++] = ys;
    if (ys >= H) jd[njd++] = ys - P;
    if (ys <= H) jn[njn++] = -ys;
    if (ys >= H && P - ys <= H && P - ys != -ys) jn[njn++] = P - ys;
    if (zs <= H) kd[nkd++] = zs;
    if (zs >= H) kd[nkd++] = zs - P;
    if (zs <= H) kn[nkn++] = -zs;
    if (zs >= H && P - zs <= H && P - zs != -zs) kn[nkn++] = P - zs;
    if (ys == 0 && zs == 0 && (xs == 0 || xs == H)) {
        float r = 0, i = 0, w = 0;
        gather_cell_all(S, t3dp, nsym, xs, 0, 0, c, weight, etab, rtab, r, i, w);
        if (xs == H) gather_cell_all(S, t3dp, nsym, -H, 0, 0, c, weight, etab, rtab, r, i, w);
        ore = r;
        ow = w;
        return;
    }
    for (int a = 0; a < njd; a++)
        for (int b = 0; b < nkd; b++) {  // direct: value and weight
            float r = 0, i = 0, w = 0;
            gather_cell_all(S, t3dp, nsym, xs, jd[a], kd[b], c, weight, etab, rtab, r, i, w);
            ore += r;
            oim += i;
            ow += w;
        }
    for (int a = 0; a < njn; a++)
        for (int b = 0; b < nkn; b++) {
            float r = 0, i = 0, w = 0;
            if (xs == 0 || xs == H) {
                // conjugate copy of the double insertion: value only
                gather_cell_all(S, t3dp, nsym, xs, jn[a], kn[b], c, weight, etab, rtab, r, i, w);
                ore += r;
                oim -= i;
                if (xs == H) {  // x = -P/2 cells: conjugate added twice, weight once
                    float r2 = 0, i2 = 0, w2 = 0;
                    gather_cell_all(S, t3dp, nsym, -H, jn[a], kn[b], c, weight, etab, rtab, r2, i2, w2);
                    ore += 2.f * r2;
                    oim -= 2.f * i2;
                    ow += w2;
                }
            } else {
                gather_cell_all(S, t3dp, nsym, -xs, jn[a], kn[b], c, weight, etab, rtab, r, i, w);
                ore += r;
                oim -= i;
                ow += w;
            }
        }
}

// Tiles of compare pixels: each workgroup takes a 32 x 16 tile of one
// (particle, symmetry op) slice and stages the slice around it (2-pixel
// halo: the sources of a compare pixel's nearest cell lie within 2 pixels
// of it) in LDS -- pixel positions, values and the insertion loop's pixel
// filters -- so that the own-contribution gather of the slice's own plane
// reads LDS; the other symmetry planes, the conjugate cells and windows
// that leave the patch take the global gather.  Same arithmetic and order
// as gather_cell.
constexpr int AT_X = 32, AT_Y = 16, AT_HALO = 2;
constexpr int AT_PX = AT_X + 2 * AT_HALO, AT_PY = AT_Y + 2 * AT_HALO;
constexpr int AT_PATCH = AT_PX * AT_PY;
// Mirror patch of the x = 0 tiles: pixels x in [-4, 2], y in
// [-ty0 - AT_Y - 1, -ty0 + 2], the sources of the mirrored cells of compare
// pixels x < 3 (around (-x, -y)).
constexpr int AT_MX = 7, AT_MY = AT_Y + 4, AT_MPATCH = AT_MX * AT_MY;

// Patch entries carry the floors of their position packed as
// (x0 + 512) | (y0 + 512) << 10 | (z0 + 512) << 20 in the .w bits, or
// AP_NOKEY for pixels the insertion loop skips (radius cut, origin range).
// For a cell key kc, kc - key has every 10-bit digit in {0, 1} exactly when
// the cell is a corner of the pixel's 2x2x2 stencil; the digits are
// differences of at most a few units between neighbouring pixels, so the
// packed difference is unambiguous, and AP_NOKEY makes it negative.
constexpr int AP_NOKEY = (int)0x80000000u;
constexpr unsigned AP_DIGITS = 0x100401u;
__device__ __forceinline__ int ap_key(int x0, int y0, int z0)
{
    return (x0 + 512) | ((y0 + 512) << 10) | ((z0 + 512) << 20);
}

struct AgreePatch {
    const float4 *pos;   // xx, yy, zz, packed floors (AP_NOKEY: skipped by the insertion loop)
    const float2 *val;   // slice_value
    int x0, y0;          // pixel of patch entry 0
};

__device__ __forceinline__ void gather_cell_patch(const float2 *__restrict__ S, const float *m, const AgreePatch &pt,
                                                  int ci, int cj, int ck, const InsertConst &c, float weight,
                                                  const float *etab, const float *rtab, float &re, float &im,
                                                  float &w)
{
    const int H = c.H;
    const float n0 = m[1] * m[5] - m[2] * m[4], n1 = m[2] * m[3] - m[0] * m[5], n2 = m[0] * m[4] - m[1] * m[3];
    const float dist = n0 * ci + n1 * cj + n2 * ck;
    if (fabsf(dist) > 1.7330f) return;  // |p - c|_inf < 1 on the plane => |n.c| < sqrt(3)
    const float s = m[0] * ci + m[1] * cj + m[2] * ck;
    const float t = m[3] * ci + m[4] * cj + m[5] * ck;
    const int xa = max(0, (int)ceilf(s - 1.7330f)), xb = min(H, (int)floorf(s + 1.7330f));
    const int ya = max(-H, (int)ceilf(t - 1.7330f)), yb = min(H - 1, (int)floorf(t + 1.7330f));
    if (xa < pt.x0 || xb >= pt.x0 + AT_PX || ya < pt.y0 || yb >= pt.y0 + AT_PY) {
        gather_cell(S, m, ci, cj, ck, c, weight, etab, rtab, re, im, w);
        return;
    }
    for (int y = ya; y <= yb; y++)
        for (int x = xa; x <= xb; x++) {
            const int e = (y - pt.y0) * AT_PX + (x - pt.x0);
            const float4 q = pt.pos[e];
            if (__float_as_int(q.w) == AP_NOKEY) continue;  // radius cut / origin out of range
            const int x0 = (int)floorf(q.x), y0 = (int)floorf(q.y), z0 = (int)floorf(q.z);
            if ((unsigned)(ci - x0) > 1u || (unsigned)(cj - y0) > 1u || (unsigned)(ck - z0) > 1u) continue;
            const float gg = corner_weight(c.mode, ci, cj, ck, q.x, q.y, q.z, c.h, weight, etab);
            const float2 dt = pt.val[e];
            re += dt.x * gg;
            im += dt.y * gg;
            w += gg;
        }
}

// Own-plane contributions to unfolded cell (ci, cj, ck), the nearest cell of
// patch entry e (row pitch PITCH), or its mirror -(ci, cj, ck) gathered
// around the mirrored pixel.  A stencil that holds the cell starts within 1 of it on every
// axis, and the cell lies within 0.5 of entry e's position, so the source
// pixel sits less than 1.5 from that position on every axis: only the
// in-plane offsets (ox, oy) in [-2, 2]^2 flagged in omask (bit
// (oy + 2) * 5 + ox + 2, uniform over the slice) can hold it.  Same terms, in
// the same y-then-x order, as gather_cell_patch.
template <int PITCH>
__device__ __forceinline__ void gather_own_offsets(const float4 *pos, const float2 *val, int e, unsigned omask, int ci,
                                                   int cj, int ck, const InsertConst &c, float weight,
                                                   const float *etab, float &re, float &im, float &w)
{
    const int kc = ap_key(ci, cj, ck);
    for (unsigned mm = omask; mm != 0u; mm &= mm - 1u) {  // set bits in increasing order
        const int k = __builtin_ctz(mm);
        const int o = e + (k / 5 - 2) * PITCH + (k % 5 - 2);
        // position, key and value read together; the weight is corner_weight()'s
        // arithmetic with the table index clamped for the misses it is computed for
        const float4 q = pos[o];
        const float2 dt = val[o];
        const bool hit = ((unsigned)(kc - __float_as_int(q.w)) & ~AP_DIGITS) == 0u;
        const float dx = (float)ci - q.x, dy = (float)cj - q.y, dz = (float)ck - q.z;
        float gg;
        if (c.mode == 0) {
            const float r = dx * dx + dy * dy + dz * dz;
            const float f = r * c.h;
            gg = etab[min((int)__builtin_fma((double)f, 50.0, 0.5), EXPN - 1)] * weight;
        } else {
            gg = (float)((1.0 - (double)fabsf(dx)) * (1.0 - (double)fabsf(dy)) * (1.0 - (double)fabsf(dz)) *
                         (double)weight);
        }
        if (hit) {
            re += dt.x * gg;
            im += dt.y * gg;
            w += gg;
        }
    }
}

// Offsets (ox, oy) in [-2, 2]^2 whose step ox * row0 + oy * row1 of slice
// rows m stays under 1.5 (plus rounding slack) on every axis: lane k < 25
// tests offset k, the wave's ballot is the mask (the whole wave active).
__device__ __forceinline__ unsigned own_offset_mask(const float *m)
{
    const int k = threadIdx.x & 63;
    bool ok = false;
    if (k < 25) {
        const float ox = (float)(k % 5 - 2), oy = (float)(k / 5 - 2);
        const float d0 = ox * m[0] + oy * m[3], d1 = ox * m[1] + oy * m[4], d2 = ox * m[2] + oy * m[5];
        ok = fmaxf(fabsf(d0), fmaxf(fabsf(d1), fabsf(d2))) < 1.51f;
    }
    return (unsigned)__ballot(ok);
}

// own_contribution() with the slice's own plane gathered from the patch.
__device__ void own_contribution_tile(const float2 *S, const float *t3dp, int nsym, int own, const AgreePatch &pt,
                                      int xs, int ys, int zs, const InsertConst &c, float weight, const float *etab,
                                      const float *rtab, float &ore, float &oim, float &ow)
{
    const int H = c.H, P = c.P;
    auto cell_all = [&](int ci, int cj, int ck, float &r, float &i, float &w) {
        for (int sop = 0; sop < nsym; sop++) {
            if (sop == own) gather_cell_patch(S, t3dp + 9 * sop, pt, ci, cj, ck, c, weight, etab, rtab, r, i, w);
            else gather_cell(S, t3dp + 9 * sop, ci, cj, ck, c, weight, etab, rtab, r, i, w);
        }
    };
    ore = oim = ow = 0.f;
    int jd[2], njd = 0, jn[2], njn = 0, kd[2], nkd = 0, kn[2], nkn = 0;
    if (ys <= H) jd[njd++] = ys;
    if (ys >= H) jd[njd++] = ys - P;
    if (ys <= H) jn[njn++] = -ys;
    if (ys >= H && P - ys <= H && P - ys != -ys) jn[njn++] = P - ys;
    if (zs <= H) kd[nkd++] = zs;
    if (zs >= H) kd[nkd++] = zs - P;
    if (zs <= H) kn[nkn++] = -zs;
    if (zs >= H && P - zs <= H && P - zs != -zs) kn[nkn++] = P - zs;
    if (ys == 0 && zs == 0 && (xs == 0 || xs == H)) {
        float r = 0, i = 0, w = 0;
        cell_all(xs, 0, 0, r, i, w);
        if (xs == H) cell_all(-H, 0, 0, r, i, w);
        ore = r;
        ow = w;
        return;
    }
    for (int a = 0; a < njd; a++)
        for (int b = 0; b < nkd; b++) {  // direct: value and weight
            float r = 0, i = 0, w = 0;
            cell_all(xs, jd[a], kd[b], r, i, w);
            ore += r;
            oim += i;
            ow += w;
        }
    for (int a = 0; a < njn; a++)
        for (int b = 0; b < nkn; b++) {
            float r = 0, i = 0, w = 0;
            if (xs == 0 || xs == H) {
                // conjugate copy of the double insertion: value only
                cell_all(xs, jn[a], kn[b], r, i, w);
                ore += r;
                oim -= i;
                if (xs == H) {  // x = -P/2 cells: conjugate added twice, weight once
                    float r2 = 0, i2 = 0, w2 = 0;
                    cell_all(-H, jn[a], kn[b], r2, i2, w2);
                    ore += 2.f * r2;
                    oim -= 2.f * i2;
                    ow += w2;
                }
            } else {
                cell_all(-xs, jn[a], kn[b], r, i, w);
                ore += r;
                oim -= i;
                ow += w;
            }
        }
}

// One workgroup per (32 x 16 compare-pixel tile, particle x symmetry op
// slice); eight double sums per particle accumulated with float64 atomics
// into acc[p][8].
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_agree_batch(const float2 *__restrict__ slices, int nsym,
                                                     const float *__restrict__ t3d, const float *__restrict__ wq,
                                                     InsertConst c, float rmax, int sub,
                                                     const float2 *__restrict__ data, const float *__restrict__ norm,
                                                     const float *__restrict__ exptab, const float *__restrict__ rtab_g,
                                                     double *__restrict__ acc_out)
{
    const int H = c.H, P = c.P;
    const int q = blockIdx.y, p = q / nsym, own = q - p * nsym;
    const int W = H + 2;
    const int ntx = (W + AT_X - 1) / AT_X;
    const int tx0 = (blockIdx.x % ntx) * AT_X, ty0 = (int)(blockIdx.x / ntx) * AT_Y - H;
    const float2 *S = slices + (size_t)p * P * (H + 1);
    const float *t3dp = t3d + 9 * (size_t)p * nsym;
    const float *m = t3d + 9 * (size_t)q;
    const float weight = wq[q];
    // LDS: patch positions and values, mirror patch positions and values,
    // fast_exp and fl(i/P) tables
    extern __shared__ float4 agree_lds[];
    float4 *ppos = agree_lds;
    float2 *pval = reinterpret_cast<float2 *>(ppos + AT_PATCH);
    float *etab = reinterpret_cast<float *>(ppos + AT_PATCH + AT_PATCH / 2 + AT_MPATCH + AT_MPATCH / 2);
    float *rtab = etab + EXPN;
    for (int i = threadIdx.x; i < EXPN; i += blockDim.x) etab[i] = exptab[i];
    for (int i = threadIdx.x; i < P + 2; i += blockDim.x) rtab[i] = rtab_g[i];
    __syncthreads();
    AgreePatch pt{ppos, pval, tx0 - AT_HALO, ty0 - AT_HALO};
    __shared__ unsigned short dq[AT_X * AT_Y];  // queued compare pixels (tile index)
    __shared__ int dn;
    if (threadIdx.x == 0) dn = 0;
    // packed floor keys need |x0|, |y0|, |z0| <= 511
    const bool fastkeys = sub && H <= 510;
    const unsigned omask = fastkeys ? own_offset_mask(m) : 0u;
    auto stage = [&](int x, int y, float4 &pos, float2 &val) {
        pos = make_float4(0.f, 0.f, 0.f, __int_as_float(AP_NOKEY));
        val = make_float2(0.f, 0.f);
        // the insertion loop's pixels: x in [0, P/2], y in [-P/2, P/2), radius cut (corners off here)
        if (x >= 0 && x <= H && y >= -H && y <= H - 1 && x * x + y * y <= H * H + H) {
            float xx, yy, zz;
            pixel_coords(x, y, m, rtab, H, c.Pf, xx, yy, zz);
            const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
            const bool in = !(x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H);
            pos = make_float4(xx, yy, zz, __int_as_float(in ? ap_key(x0, y0, z0) : AP_NOKEY));
            if (in) val = slice_value(S, x, y, P, H);
        }
    };
    for (int e = threadIdx.x; e < AT_PATCH; e += blockDim.x)
        stage(pt.x0 + e % AT_PX, pt.y0 + e / AT_PX, ppos[e], pval[e]);
    float4 *mpos = ppos + AT_PATCH + AT_PATCH / 2;  // after pval
    float2 *mval = reinterpret_cast<float2 *>(mpos + AT_MPATCH);
    const int my0 = -ty0 - AT_Y - 1;
    const bool mirror = fastkeys && tx0 == 0;
    if (mirror)
        for (int e = threadIdx.x; e < AT_MPATCH; e += blockDim.x)
            stage(e % AT_MX - 4, my0 + e / AT_MX, mpos[e], mval[e]);
    __syncthreads();
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // One compare pixel.  Outside the general pass a pixel whose voxel needs
    // the general own-contribution enumeration returns false untouched, to be
    // queued: the few such pixels of a tile (x < 3, voxels on the x = 0,
    // x = P/2 and y, z = P/2 planes) then share waves instead of holding back
    // every wave they fall in.
    auto pixel = [&](int x, int y, bool general) -> bool {
        const float rx = rtab[x + H], ry = rtab[y + H];
        const float c0 = rx * m[0] + ry * m[3];
        const float c1 = rx * m[1] + ry * m[4];
        const float c2 = rx * m[2] + ry * m[5];
        if (fabsf(c0) > 0.5f || fabsf(c1) >= 0.5f || fabsf(c2) >= 0.5f) return true;
        const float xx = c0 * c.Pf, yy = c1 * c.Pf, zz = c2 * c.Pf;
        const int x0 = round_ref(xx), y0 = round_ref(yy), z0 = round_ref(zz);
        // Common voxel: |x0|, |y0|, |z0| < P/2 leave one direct and one
        // mirrored source cell, +-(x0, y0, z0) (on x0 = 0 the mirrored one
        // through the double insertion's conjugate copy, value only).  From
        // x >= 3 on, the own plane (pixels x >= 0) cannot reach the mirrored
        // cell, whose in-plane distance exceeds sqrt(3) + sqrt(3)/2; below,
        // its sources lie around (-x, -y), in the mirror patch.
        const bool common = fastkeys && (x >= 3 || mirror) && abs(x0) < H && abs(y0) < H && abs(z0) < H;
        if (sub && !general && !common) return false;
        const float *Sf = reinterpret_cast<const float *>(S);
        const int slen = (P + 2) * P;
        const int fidx = 2 * x + (P + 2) * (y < 0 ? P + y : y);
        const float d20 = fidx < slen ? Sf[fidx] : 0.f;
        const float d21 = fidx + 1 < slen ? Sf[fidx + 1] : 0.f;
        int xs, ysn, zsn;
        if (x0 < 0) {
            xs = -x0;
            ysn = y0 <= 0 ? -y0 : P - y0;
            zsn = z0 <= 0 ? -z0 : P - z0;
        } else {
            xs = x0;
            ysn = y0 < 0 ? P + y0 : y0;
            zsn = z0 < 0 ? P + z0 : z0;
        }
        const size_t ci = cidx(xs, ysn, zsn, P, H);
        float vr = 0.f, vi = 0.f;
        const bool inb = !(abs(x0) >= H + 1 || abs(y0) > H || abs(z0) > H);
        if (inb) {
            const float2 v = data[ci];
            vr = v.x;
            vi = v.y;
        }
        float nv = norm[ci];
        if (sub) {
            float ore, oim, ow;
            if (general) {
                own_contribution_tile(S, t3dp, nsym, own, pt, xs, ysn, zsn, c, weight, etab, rtab, ore, oim, ow);
            } else {
                const int e = (y - pt.y0) * AT_PX + (x - pt.x0);
                float r = 0.f, i = 0.f, w = 0.f, r2 = 0.f, i2 = 0.f, w2 = 0.f;
                for (int sop = 0; sop < nsym; sop++) {
                    if (sop == own) {
                        gather_own_offsets<AT_PX>(ppos, pval, e, omask, x0, y0, z0, c, weight, etab, r, i, w);
                        if (x < 3)  // mirror patch entry of (-x, -y)
                            gather_own_offsets<AT_MX>(mpos, mval, 4 - x + (-y - my0) * AT_MX, omask, -x0, -y0, -z0,
                                                      c, weight, etab, r2, i2, w2);
                    } else {
                        gather_cell(S, t3dp + 9 * sop, x0, y0, z0, c, weight, etab, rtab, r, i, w);
                        gather_cell(S, t3dp + 9 * sop, -x0, -y0, -z0, c, weight, etab, rtab, r2, i2, w2);
                    }
                }
                if (x0 >= 0) {  // direct cell (x0, y0, z0), mirrored -(x0, y0, z0)
                    ore = r + r2;
                    oim = i - i2;
                    ow = x0 == 0 ? w : w + w2;
                } else {
                    ore = r2 + r;
                    oim = i2 - i;
                    ow = w2 + w;
                }
            }
            if (inb) {
                vr = vr - ore;
                vi = vi - oim;
            }
            nv = nv - ow;
        }
        if (x0 < 0) vi = -vi;  // get_complex_at conjugate for x<0
        if (nv != 0.f) {
            const float dt0 = __fdiv_rn(vr, nv), dt1 = __fdiv_rn(vi, nv), dt2 = nv;
            const float r = sqrt_int_rn(x * x + y * y);  // Util::hypot_fast
            const float dotv = (dt0 * d20 + dt1 * d21) * dt2;
            const float pw = (dt0 * dt0 + dt1 * dt1) * dt2;
            const float pw2 = (d20 * d20 + d21 * d21) * dt2;
            if (r > 3 && r < rmax) {
                acc[4] += dotv;
                acc[5] += pw;
                acc[6] += pw2;
            }
            acc[1] += dt2;
            acc[0] += dotv;
            acc[2] += pw;
            acc[3] += pw2;
            acc[7] += 1.0;
        }
        return true;
    };
    // tile slot l -> pixel; in the x = 0 tiles the first 48 slots are the
    // columns x < 3 of all 16 rows, so that only one wave runs their mirrored
    // gather
    auto slot = [&](int l, int &x, int &y) {
        if (tx0 == 0) {
            if (l < 3 * AT_Y) {
                x = l % 3;
                y = ty0 + l / 3;
            } else {
                x = 3 + (l - 3 * AT_Y) % (AT_X - 3);
                y = ty0 + (l - 3 * AT_Y) / (AT_X - 3);
            }
        } else {
            x = tx0 + l % AT_X;
            y = ty0 + l / AT_X;
        }
    };
    for (int it = 0; it < AT_Y * AT_X / 256; it++) {
        const int l = threadIdx.x + it * 256;
        int x, y;
        slot(l, x, y);
        if (x >= W || y >= H || (x == 0 && y == 0)) continue;
        if (!pixel(x, y, false)) dq[atomicAdd(&dn, 1)] = (unsigned short)l;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < dn; j += blockDim.x) {
        int x, y;
        slot(dq[j], x, y);
        pixel(x, y, true);
    }
    // block sums through LDS (the patch's space, no longer read): 32 threads
    // per sum add 8 of the 256 partials each, then a 32-lane butterfly
    __syncthreads();
    static_assert((sizeof(float4) + sizeof(float2)) * AT_PATCH >= 8 * 256 * sizeof(double), "block sums");
    double *red = reinterpret_cast<double *>(agree_lds);  // 8 x 256 doubles within the patch's bytes
#pragma unroll
    for (int k = 0; k < 8; k++) red[k * 256 + threadIdx.x] = acc[k];
    __syncthreads();
    const int k = threadIdx.x >> 5, l = threadIdx.x & 31;
    double v = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) v += red[k * 256 + l + 32 * i];
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (l == 0) unsafeAtomicAdd(&acc_out[(size_t)p * 8 + k], v);
}

hipError_t launch_agree_batch(const float2 *slices, int nsym, const float *t3d, const float *wq, int nslices,
                              InsertConst c, float rmax, int sub, const float2 *data, const float *norm,
                              const float *exptab, const float *rtab, double *acc, hipStream_t s)
{
    if (nslices <= 0) return hipSuccess;
    c.corners = 0;
    const int ntx = (c.P / 2 + 2 + AT_X - 1) / AT_X, nty = (c.P + AT_Y - 1) / AT_Y;
    const size_t lds = (sizeof(float4) + sizeof(float2)) * (AT_PATCH + AT_MPATCH) + sizeof(float) * (EXPN + c.P + 2);
    hipLaunchKernelGGL(k_agree_batch, dim3(ntx * nty, nslices), dim3(256), lds, s, slices, nsym, t3d, wq, c, rmax, sub,
                       data, norm, exptab, rtab, acc);
    return hipGetLastError();
}

// ------------------------------------------------------------ projection
// FourierReconstructor::projection (libEM/reconstructor.cpp:1383-1445): the
// slice pixel (x, y) takes the nearest voxel's value over its weight
// (pixel_at, :1448-1461) when the weight is at least 0.005, halved on the
// x = 0 column.  The result is what set_complex_at (emdata_core.cpp:171-205)
// leaves in the (P+2) x P image after the reference's loop: on the x = 0 and
// x = P/2 columns each pixel also writes its conjugate at -y, so a stored
// row holds the later of the two iterations that write it (y > 0 comes after
// y < 0), rows 0 and P/2 keep the real part only.
__device__ __forceinline__ bool project_pixel(int x, int y, const Mat9 &m, const InsertConst &c,
                                              const float *__restrict__ rtab, const float2 *__restrict__ data,
                                              const float *__restrict__ norm, float &re, float &im)
{
    const int H = c.H, P = c.P;
    const float rx = rtab[x + H], ry = rtab[y + H];
    const float c0 = rx * m.m[0] + ry * m.m[3];
    const float c1 = rx * m.m[1] + ry * m.m[4];
    const float c2 = rx * m.m[2] + ry * m.m[5];
    if (fabsf(c0) > 0.5f || fabsf(c1) >= 0.5f || fabsf(c2) >= 0.5f) return false;
    const int x0 = round_ref(c0 * c.Pf), y0 = round_ref(c1 * c.Pf), z0 = round_ref(c2 * c.Pf);
    size_t ci;
    if (x0 < 0) ci = cidx(-x0, y0 <= 0 ? -y0 : P - y0, z0 <= 0 ? -z0 : P - z0, P, H);
    else ci = cidx(x0, y0 < 0 ? P + y0 : y0, z0 < 0 ? P + z0 : z0, P, H);
    const float nv = norm[ci];
    if ((double)nv < 0.005) return false;  // dt[2] < 0.005, a double comparison
    const float2 v = data[ci];
    re = __fdiv_rn(v.x, nv);
    im = __fdiv_rn(x0 < 0 ? -v.y : v.y, nv);
    if (x == 0 && y != 0 && y != -H) {
        re /= 2;
        im /= 2;
    }
    return true;
}

// shift: the translation's phase ramp of the "xform" processor's complex
// branch (libEM/processor.cpp:11081-11176; with no rotation its bilinear
// weights are exactly 1, 0, 0, 0), the reference's long double pi
// arithmetic restated in double.  center: xform.phaseorigin.tocenter on the
// complex image, x (-1)^(x + row) (sparx/emdata_sparx.cpp:8451-8460).
__global__ void __launch_bounds__(256) k_project(const float2 *__restrict__ data, const float *__restrict__ norm,
                                                 InsertConst c, Mat9 m, const float *__restrict__ rtab, float xshift,
                                                 float yshift, int shift, int center, float2 *__restrict__ out)
{
    const int H = c.H, P = c.P;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * (H + 1)) return;
    const int x = idx % (H + 1), r = idx / (H + 1);
    float re = 0.f, im = 0.f, pr, pim;
    if (x != 0 && x != H) {
        if (!project_pixel(x, r < H ? r : r - P, m, c, rtab, data, norm, re, im)) re = im = 0.f;
    } else if (r == 0 || r == H) {
        if (project_pixel(x, r == 0 ? 0 : -H, m, c, rtab, data, norm, pr, pim)) re = pr;
    } else {
        const int a = r < H ? r : P - r;
        if (project_pixel(x, a, m, c, rtab, data, norm, pr, pim)) {
            re = pr;
            im = r < H ? pim : -pim;
        } else if (project_pixel(x, -a, m, c, rtab, data, norm, pr, pim)) {
            re = pr;
            im = r < H ? -pim : pim;
        }
    }
    if (shift) {
        const double pi = 3.141592653589793238462643383279502884;
        const int ky = r >= H + 1 ? r - P : r;
        const float phaseY = (float)(-2 * pi * ky * (double)yshift / P);
        const float phase = (float)((double)phaseY - 2 * pi * x * (double)xshift / P);
        const double cp = cos((double)phase), sp = sin((double)phase);
        const float rb = (float)((double)re * cp - (double)im * sp);
        const float ib = (float)((double)re * sp + (double)im * cp);
        re = rb;
        im = ib;
    }
    if (center && ((x + r) & 1)) {
        re = -re;
        im = -im;
    }
    out[idx] = make_float2(re, im);
}

hipError_t launch_project(const float2 *data, const float *norm, InsertConst c, const float *rot9, const float *rtab,
                          float xshift, float yshift, int center, float2 *out, hipStream_t s)
{
    Mat9 m;
    for (int i = 0; i < 9; i++) m.m[i] = rot9[i];
    const int n = c.P * (c.H + 1);
    const int shift = xshift != 0.0f || yshift != 0.0f;
    hipLaunchKernelGGL(k_project, dim3((n + 255) / 256), dim3(256), 0, s, data, norm, c, m, rtab, xshift, yshift, shift,
                       center, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ gauss_fft projector
// GaussFFTProjector::project3d (libEM/projector.cpp:69-212) and interp_ft_3d
// (:216-295) for modes 0-2 on the map's transform kept in the volume layout
// (phase origin at the corner, rows and planes in wrap order).  The
// reference interpolates the complex volume with its Fourier origin moved
// to the centre in y and z and reads it by linear float index
// 2 x + y (P+2) + z (P+2) P; pv_lin maps such an index (0 outside the array)
// back to the wrap-ordered storage.  Per projection q: rec[12 q ..] =
// inverse rotation rows (9), scale, tx, ty.  Output element (x, r) of
// projection q is the slice pixel (x, row (r + P/2) mod P of the centred
// slice): the Fourier origin back in the corner.
__device__ __forceinline__ float pv_lin(const float2 *__restrict__ pv, long long L, int P, int H)
{
    const int nx = P + 2;
    if (L < 0 || L >= (long long)nx * P * P) return 0.f;
    const int xr = (int)(L % nx);
    const long long q = L / nx;
    const int y = (int)(q % P), z = (int)(q / P);
    const float2 v = pv[cidx(xr >> 1, y + H < P ? y + H : y + H - P, z + H < P ? z + H : z + H - P, P, H)];
    return (xr & 1) ? v.y : v.x;
}

__device__ __forceinline__ float agauss_f(float a, float dx, float dy, float dz, float d)
{
    return a * expf(-(dx * dx + dy * dy + dz * dz) / d);  // Util::agauss (util.h:912-915)
}

__device__ bool gfft_interp(int mode, const float2 *__restrict__ pv, int P, int H, float x, float y, float z,
                            float &d0, float &d1, float gw)
{
    const int nx = P + 2, ny = P, nz = P;
    if (mode == 0) {
        const int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
        float a0 = 0.f, a1 = 0.f, nrm = 0.f;
        for (int ix = x0; ix <= x0 + 1; ix++)
            for (int iy = y0; iy <= y0 + 1; iy++)
                for (int iz = z0; iz <= z0 + 1; iz++) {
                    const float w = (float)((1.0 - (double)fabsf((float)ix - x)) * (1.0 - (double)fabsf((float)iy - y)) *
                                            (1.0 - (double)fabsf((float)iz - z)));
                    const long long L = 2LL * ix + (long long)iy * nx + (long long)iz * nx * ny;
                    a0 += w * pv_lin(pv, L, P, H);
                    a1 += w * pv_lin(pv, L + 1, P, H);
                    nrm += w;
                }
        d0 = a0 / nrm;
        d1 = a1 / nrm;
        return true;
    }
    if (mode == 1) {
        const int x0 = 2 * (int)floor(x + 0.5f), y0 = (int)floor(y + 0.5f), z0 = (int)floor(z + 0.5f);
        const long long L = (long long)x0 + (long long)y0 * nx + (long long)z0 * nx * ny;
        d0 = pv_lin(pv, L, P, H);
        d1 = pv_lin(pv, L + 1, P, H);
        return true;
    }
    const int x0 = (int)floor(x), y0 = (int)floor(y), z0 = (int)floor(z);
    const float dx = x - x0, dy = y - y0, dz = z - z0;
    if (!(x0 <= nx / 2 - 2 && y0 <= ny - 1 && z0 <= nz - 1)) return false;
    const long long i0 = (long long)x0 * 2 + (long long)y0 * nx + (long long)z0 * nx * ny;
    const float n = agauss_f(1, dx, dy, dz, gw) + agauss_f(1, 1 - dx, dy, dz, gw) + agauss_f(1, dx, 1 - dy, dz, gw) +
                    agauss_f(1, 1 - dx, 1 - dy, dz, gw) + agauss_f(1, dx, dy, 1 - dz, gw) +
                    agauss_f(1, 1 - dx, dy, 1 - dz, gw) + agauss_f(1, dx, 1 - dy, 1 - dz, gw) +
                    agauss_f(1, 1 - dx, 1 - dy, 1 - dz, gw);
    float d[2];
    for (int c = 0; c < 2; c++) {
        const long long i = i0 + c;
        const long long nxy = (long long)nx * ny;
        // the reference divides its last term only by n (operator precedence)
        d[c] = agauss_f(pv_lin(pv, i, P, H), dx, dy, dz, gw) + agauss_f(pv_lin(pv, i + 2, P, H), 1 - dx, dy, dz, gw) +
               agauss_f(pv_lin(pv, i + nx, P, H), dx, 1 - dy, dz, gw) +
               agauss_f(pv_lin(pv, i + nx + 2, P, H), 1 - dx, 1 - dy, dz, gw) +
               agauss_f(pv_lin(pv, i + nxy, P, H), dx, dy, 1 - dz, gw) +
               agauss_f(pv_lin(pv, i + 2 + nxy, P, H), 1 - dx, dy, 1 - dz, gw) +
               agauss_f(pv_lin(pv, i + nx + nxy, P, H), dx, 1 - dy, 1 - dz, gw) +
               agauss_f(pv_lin(pv, i + 2 + nx + nxy, P, H), 1 - dx, 1 - dy, 1 - dz, gw) / n;
    }
    d0 = d[0];
    d1 = d[1];
    return true;
}

// ramp: the translation's phase ramp (the "xform" processor's complex
// branch, as in k_project), for the Fourier-slice output
__global__ void __launch_bounds__(256) k_gauss_fft(const float2 *__restrict__ pv, int P, int mode, float gw,
                                                   const float *__restrict__ rec, int ramp, float2 *__restrict__ out)
{
    const int H = P / 2;
    const int q = blockIdx.y;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * (H + 1)) return;
    const int x = idx % (H + 1), r = idx / (H + 1);
    const int yc = r + H < P ? r + H : r + H - P;  // the centred slice's row
    const float *m = rec + 12 * (size_t)q;
    float re = 0.f, im = 0.f;
    const int yy_i = yc - P / 2;
    if ((long long)x * x + (long long)yy_i * yy_i < (long long)(P / 2 - 1) * (P / 2 - 1)) {
        const float cx = (float)x, cy = (float)yy_i, cz = 0.0f;
        float xx = m[0] * cx + m[1] * cy + m[2] * cz + 0.0f;
        float yy = m[3] * cx + m[4] * cy + m[5] * cz + 0.0f;
        float zz = m[6] * cx + m[7] * cy + m[8] * cz + 0.0f;
        int cc = 1;
        if (xx < 0) {
            xx = -xx;
            yy = -yy;
            zz = -zz;
            cc = -1;
        }
        const float scale = m[9];
        if (scale != 1.0f) {
            xx *= scale;
            yy *= scale;
            zz *= scale;
        }
        yy += (float)(P / 2);
        zz += (float)(P / 2);
        if (!(yy < 0 || xx < 0 || zz < 0)) {
            float d0, d1;
            if (gfft_interp(mode, pv, P, H, xx, yy, zz, d0, d1, gw)) {
                re = d0;
                im = d1 * (float)cc;
            }
        }
    }
    const float xshift = m[10], yshift = m[11];
    if (ramp && (xshift != 0.0f || yshift != 0.0f)) {
        const double pi = 3.141592653589793238462643383279502884;
        const int ky = r >= H + 1 ? r - P : r;
        const float phaseY = (float)(-2 * pi * ky * (double)yshift / P);
        const float phase = (float)((double)phaseY - 2 * pi * x * (double)xshift / P);
        const double cp = cos((double)phase), sp = sin((double)phase);
        const float rb = (float)((double)re * cp - (double)im * sp);
        const float ib = (float)((double)re * sp + (double)im * cp);
        re = rb;
        im = ib;
    }
    out[(size_t)q * P * (H + 1) + idx] = make_float2(re, im);
}

// math.gausskernelfix on n P x P images (MaskGaussInvProcessor,
// processor.h:4131-4176): each pixel over exp(-dist sv), dist the squared
// distance from (P/2, P/2), sv = gauss_width / P^2
__global__ void k_gausskernelfix2d(float *__restrict__ img, int P, int n, float sv)
{
    const size_t tot = (size_t)n * P * P;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % P), y = (int)((i / P) % P);
        const float xc = (float)(P / 2), yc = (float)(P / 2);
        const float dist = ((float)x - xc) * ((float)x - xc) + ((float)y - yc) * ((float)y - yc) + 0.0f;
        img[i] = img[i] / expf(-dist * sv);
    }
}

hipError_t launch_gauss_fft(const float2 *pv, int P, int n, int mode, float gw, const float *rec, int ramp,
                            float2 *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int m = P * (P / 2 + 1);
    hipLaunchKernelGGL(k_gauss_fft, dim3((m + 255) / 256, n), dim3(256), 0, s, pv, P, mode, gw, rec, ramp, out);
    return hipGetLastError();
}

hipError_t launch_gausskernelfix2d(float *img, int P, int n, float gauss_width, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const float sv = gauss_width / (float)(P * P);
    const unsigned blocks = (unsigned)std::min<size_t>(((size_t)n * P * P + 255) / 256, 65536);
    hipLaunchKernelGGL(k_gausskernelfix2d, dim3(blocks), dim3(256), 0, s, img, P, n, sv);
    return hipGetLastError();
}

hipError_t launch_agree(const float2 *slices, int nsym, const float *t3d, int nslices, InsertConst c, float rmax,
                        const float2 *data, const float *norm, const float *rtab, double *partial, hipStream_t s)
{
    if (nslices <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_agree, dim3(agree_blocks(c.P), nslices), dim3(256), 0, s, slices, nsym, t3d, c, rmax, data,
                       norm, rtab, partial);
    return hipGetLastError();
}

// ------------------------------------------------------------ binning
// Brick b = (bx, by, bz) owns the stencil origins [lo, lo + bs) per axis,
// lo = ow + bs*b: a pixel is added by the one brick holding its stencil
// origin, i.e. p in [po + bs*b, po + bs*(b+1)) on every axis (for gauss_2
// the origin is floor(p), po = ow = -H-1, and the 8 corners land in the
// brick's (bs+1)^3 cells [lo, lo+bs]^3).  Pixel positions lie on the
// plane n.p = 0 (n = u x v of the slice's t3d rows), so per brick column
// along the dominant normal axis at most 4 bricks qualify.  A slice is
// binned under key 3 * brick + (its dominant normal axis), so that every
// work item's slices share one column direction.
//
// The binning is owned by brick columns: one workgroup per (column, axis)
// runs over the slices of that dominant axis and counts / lists the
// column's bricks in LDS, so no (slice, brick) pair touches a global atomic.

// Per slice: the plane record, appended to the segment of its dominant axis
// (one atomic per wave and axis).  Weight-0 slices are skipped
// (reconstructor.cpp:1088).
__global__ void __launch_bounds__(256) k_slice_recs(const float *__restrict__ t3d, const float *__restrict__ wq,
                                                    int nslices, BinBuffers bb)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int a = -1;
    SliceRec r;
    if (q < nslices && wq[q] != 0.0f) {
        const float *m = t3d + 9 * (size_t)q;
        double u[3], v[3];
        for (int d = 0; d < 3; d++) {
            r.u[d] = m[d];
            r.v[d] = m[3 + d];
            u[d] = m[d];
            v[d] = m[3 + d];
        }
        const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double a0 = fabs(n[0]), a1 = fabs(n[1]), a2 = fabs(n[2]);
        a = (a0 >= a1 && a0 >= a2) ? 0 : (a1 >= a2 ? 1 : 2);
        const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
        r.r1 = -n[b1] / n[a];
        r.r2 = -n[b2] / n[a];
        r.q = q;
        r.pad = 0;
    }
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const unsigned long long m = __ballot(a == ax);
        if (m == 0) continue;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&bb.ctl[4 + ax], __popcll(m));
        base = __shfl(base, leader, 64);
        if (a == ax) {
            const int rank = __popcll(m & ((1ull << lane) - 1ull));
            bb.recs[(size_t)ax * bb.rec_cap + base + rank] = r;
        }
    }
}

// Does the slice's pixel domain (half-disk x>=0 of radius sqrt(H^2+H), or
// the half-square with corners) come within reach of the brick?
__device__ __forceinline__ bool brick_in_domain(const SliceRec &f, const int bc[3], int H, int BS, double po,
                                                int corners, double rr)
{
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = po + BS * bc[d] + 0.5 * BS;
    const double s = c[0] * (double)f.u[0] + c[1] * (double)f.u[1] + c[2] * (double)f.u[2];
    const double t = c[0] * (double)f.v[0] + c[1] * (double)f.v[1] + c[2] * (double)f.v[2];
    const double rh = 0.8660254037844386 * BS + 0.01;
    if (s < -rh) return false;
    if (corners) return s <= H + rh && t >= -H - rh && t <= H - 1 + rh;
    return s * s + t * t <= rr * rr;
}

// One workgroup per (brick column (c1, c2), axis a): MODE 0 counts the
// slices of each brick of the column, MODE 1 lists them from the key
// offsets on.  LDS: one counter per brick of the column.
template <int MODE>
__global__ void __launch_bounds__(256) k_bin_cols(InsertConst c, BrickGeom g, BinBuffers bb)
{
    extern __shared__ int colcnt[];
    const int nb = g.nb, BS = g.bs, H = c.H;
    const double po = g.po;
    const int col = blockIdx.x, A = blockIdx.y;
    const int c1 = col % nb, c2 = col / nb;
    const int b1 = A == 0 ? 1 : 0, b2 = A == 2 ? 1 : 2;
    int bc[3];
    bc[b1] = c1;
    bc[b2] = c2;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        bc[A] = i;
        colcnt[i] = MODE == 0 ? 0 : bb.offsets[3 * ((bc[2] * nb + bc[1]) * nb + bc[0]) + A];
    }
    __syncthreads();
    const double dlt = 1e-3;
    const double L1 = po + BS * c1 - dlt, U1 = po + BS * c1 + BS + dlt;
    const double L2 = po + BS * c2 - dlt, U2 = po + BS * c2 + BS + dlt;
    const double rr = sqrt((double)c.rcut * c.rcut + c.rcut) + 0.8660254037844386 * BS + 0.01;
    const int n = bb.ctl[4 + A];
    const SliceRec *recs = bb.recs + (size_t)A * bb.rec_cap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const SliceRec f = recs[i];
        const double pa0 = f.r1 * L1 + f.r2 * L2, pa1 = f.r1 * U1 + f.r2 * L2;
        const double pa2 = f.r1 * L1 + f.r2 * U2, pa3 = f.r1 * U1 + f.r2 * U2;
        const double amin = fmin(fmin(pa0, pa1), fmin(pa2, pa3)), amax = fmax(fmax(pa0, pa1), fmax(pa2, pa3));
        int lo = (int)floor((amin - dlt - po) / BS);
        int hi = (int)floor((amax + dlt - po) / BS);
        if (lo < 0) lo = 0;
        if (hi > nb - 1) hi = nb - 1;
        for (int ca = lo; ca <= hi; ca++) {
            bc[A] = ca;
            if (!brick_in_domain(f, bc, H, BS, po, c.corners, rr)) continue;
            const int pos = atomicAdd(&colcnt[ca], 1);
            if (MODE == 1) {
                if (pos < bb.list_cap) bb.list[pos] = f.q;
                else bb.ctl[3] = 1;
            }
        }
    }
    if (MODE == 0) {
        __syncthreads();
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            bc[A] = i;
            bb.counts[3 * ((bc[2] * nb + bc[1]) * nb + bc[0]) + A] = colcnt[i];
        }
    }
}

// Exclusive scans of the key counts (-> list offsets) and of the work items
// per key (ceil(count / K) -> item offsets) in three passes: per-tile sums,
// a scan of the tile sums with the totals, and the per-key offsets.  With
// per_brick the items of a brick span its three axis keys (whose lists are
// consecutive): key 3b holds ceil((c[3b] + c[3b+1] + c[3b+2]) / K) items,
// keys 3b+1 and 3b+2 none.
constexpr int SCAN_TILE = 1024;  // keys per tile: 256 threads x 4

__device__ __forceinline__ int key_items(const int *__restrict__ counts, int k, int nkeys, int ck, int K,
                                         int per_brick)
{
    if (!per_brick) return (ck + K - 1) / K;
    if (k % 3) return 0;
    const int c = ck + (k + 1 < nkeys ? counts[k + 1] : 0) + (k + 2 < nkeys ? counts[k + 2] : 0);
    return (c + K - 1) / K;
}

__device__ __forceinline__ int4 load_counts4(const int *counts, int k, int n)
{
    if (k + 3 < n) return *reinterpret_cast<const int4 *>(counts + k);
    int4 v;
    v.x = k < n ? counts[k] : 0;
    v.y = k + 1 < n ? counts[k + 1] : 0;
    v.z = k + 2 < n ? counts[k + 2] : 0;
    v.w = 0;
    return v;
}

// 256-thread exclusive scan of (a, b); returns the block totals
__device__ __forceinline__ void block_scan2(int &a, int &b, int &ta, int &tb)
{
    __shared__ int sa[4], sb[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int va = __shfl_up(ia, off, 64), vb = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += va;
            ib += vb;
        }
    }
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    int pa = 0, pb = 0;
    for (int w = 0; w < wave; w++) {
        pa += sa[w];
        pb += sb[w];
    }
    ta = sa[0] + sa[1] + sa[2] + sa[3];
    tb = sb[0] + sb[1] + sb[2] + sb[3];
    a = pa + ia - a;
    b = pb + ib - b;
}

__global__ void __launch_bounds__(256) k_scan_tiles(BrickGeom g, BinBuffers bb, int K)
{
    const int k = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    const int4 c = load_counts4(bb.counts, k, g.nkeys);
    const int cv[4] = {c.x, c.y, c.z, c.w};
    int a = c.x + c.y + c.z + c.w;
    int b = 0;
#pragma unroll
    for (int t = 0; t < 4; t++)
        if (k + t < g.nkeys) b += key_items(bb.counts, k + t, g.nkeys, cv[t], K, g.per_brick);
    int ta, tb;
    block_scan2(a, b, ta, tb);
    if (threadIdx.x == 0) {
        bb.tile[2 * blockIdx.x] = ta;
        bb.tile[2 * blockIdx.x + 1] = tb;
    }
}

__global__ void __launch_bounds__(1024) k_scan_top(BrickGeom g, BinBuffers bb, int ntiles)
{
    const int tid = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const int t0 = tid * per, t1 = min(ntiles, t0 + per);
    long long sc = 0, si = 0;
    for (int t = t0; t < t1; t++) {
        sc += bb.tile[2 * t];
        si += bb.tile[2 * t + 1];
    }
    __shared__ long long s_c[1024], s_i[1024];
    s_c[tid] = sc;
    s_i[tid] = si;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long vc = tid >= off ? s_c[tid - off] : 0;
        const long long vi = tid >= off ? s_i[tid - off] : 0;
        __syncthreads();
        s_c[tid] += vc;
        s_i[tid] += vi;
        __syncthreads();
    }
    long long ec = s_c[tid] - sc, ei = s_i[tid] - si;
    for (int t = t0; t < t1; t++) {
        const int a = bb.tile[2 * t], b = bb.tile[2 * t + 1];
        bb.tile[2 * t] = (int)ec;
        bb.tile[2 * t + 1] = (int)ei;
        ec += a;
        ei += b;
    }
    if (tid == 1023) {
        const long long tc = s_c[1023], ti = s_i[1023];
        bb.offsets[g.nkeys] = (int)tc;
        bb.item_off[g.nkeys] = (int)ti;
        bb.ctl[0] = 0;
        bb.ctl[1] = (int)(ti < bb.item_cap ? ti : bb.item_cap);
        bb.ctl[2] = (int)(tc < bb.list_cap ? tc : bb.list_cap);
        if (tc > bb.list_cap || ti > bb.item_cap) bb.ctl[3] = 1;
    }
}

__global__ void __launch_bounds__(256) k_scan_apply(BrickGeom g, BinBuffers bb, int K)
{
    const int k = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    const int4 c = load_counts4(bb.counts, k, g.nkeys);
    const int cv[4] = {c.x, c.y, c.z, c.w};
    int iv[4];
#pragma unroll
    for (int t = 0; t < 4; t++) iv[t] = k + t < g.nkeys ? key_items(bb.counts, k + t, g.nkeys, cv[t], K, g.per_brick) : 0;
    int a = c.x + c.y + c.z + c.w;
    int b = iv[0] + iv[1] + iv[2] + iv[3];
    int ta, tb;
    block_scan2(a, b, ta, tb);
    a += bb.tile[2 * blockIdx.x];
    b += bb.tile[2 * blockIdx.x + 1];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (k + t < g.nkeys) {
            bb.offsets[k + t] = a;
            bb.item_off[k + t] = b;
        }
        a += cv[t];
        b += iv[t];
    }
}

int scan_tiles(const BrickGeom &g) { return (g.nkeys + SCAN_TILE - 1) / SCAN_TILE; }

__global__ void k_bin_items(BrickGeom g, BinBuffers bb, int K)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= g.nkeys) return;
    if (g.per_brick && b % 3) return;  // the brick's items are listed under key 3 * brick
    const int cnt = g.per_brick ? bb.counts[b] + (b + 1 < g.nkeys ? bb.counts[b + 1] : 0) +
                                      (b + 2 < g.nkeys ? bb.counts[b + 2] : 0)
                                : bb.counts[b];
    const int base = bb.item_off[b];
    const int off = bb.offsets[b];
    for (int i = 0; i * K < cnt; i++) {
        if ((long long)base + i >= bb.item_cap) break;
        WorkItem w;
        w.key = b;
        w.start = off + i * K;
        w.count = min(K, cnt - i * K);
        if ((long long)w.start + w.count > bb.list_cap) w.count = (int)max(0LL, bb.list_cap - w.start);
        w.pad = 0;
        bb.items[base + i] = w;
    }
}

hipError_t launch_bin(const float *t3d, const float *wq, int nslices, InsertConst c, BrickGeom g, BinBuffers b,
                      int K, hipStream_t s)
{
    hipError_t e;
    if (nslices > b.rec_cap) return hipErrorInvalidValue;
    if ((e = launch_fill_i32(b.counts, 0, g.nkeys, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32(b.ctl, 0, 8, s)) != hipSuccess) return e;
    const dim3 cgrid(g.nb * g.nb, 3);
    const size_t clds = sizeof(int) * g.nb;
    if (nslices > 0) {
        hipLaunchKernelGGL(k_slice_recs, dim3((nslices + 255) / 256), dim3(256), 0, s, t3d, wq, nslices, b);
        hipLaunchKernelGGL(k_bin_cols<0>, cgrid, dim3(256), clds, s, c, g, b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int nt = scan_tiles(g);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(256), 0, s, g, b, K);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, g, b, nt);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(256), 0, s, g, b, K);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (nslices > 0) {
        hipLaunchKernelGGL(k_bin_cols<1>, cgrid, dim3(256), clds, s, c, g, b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bin_items, dim3((g.nkeys + 255) / 256), dim3(256), 0, s, g, b, K);
    return hipGetLastError();
}

// ------------------------------------------------------------ brick scatter insertion
// A work item is (brick of 8^3 stencil origins, axis A, up to K slices whose
// dominant normal axis is A).  Its wave owns a private accumulator of the
// brick's 9^3 cells in LDS and walks the slice pixels whose stencil origin
// lies in the brick in windows of 64 (one pixel per lane), adding each
// pixel's 8 corners with plain LDS read-modify-writes; every pixel is added
// by exactly one brick, and the cells on a brick's high faces are summed
// with the next bricks' in the merge:
//  - two lanes whose pixels share a stencil origin cell would race on all
//    eight corners; an LDS tag per origin cell elects one lane per cell, the
//    others stay pending for the next window (no atomics);
//  - two lanes with different origins can still meet on a cell through
//    different corners, so the corners are processed one after another with
//    a wave-level ordering point between them.
// The accumulator is three float arrays (re, im, w) over the cells in
// [A][B2][B1] order (B1 < B2 the other axes): one corner's cells of a slice
// with dominant axis A fall in distinct (B1, B2) columns, so the lanes of an
// access spread over the 32 banks.
constexpr int ACC_N = HALO_CELLS;                      // cells per field
constexpr int TAG_N = BRICK * BRICK * BRICK;           // stencil origins in [lo, lo+7]^3
constexpr int ACC_BYTES = (12 * ACC_N + 15) & ~15;     // re, im, w
constexpr int REC_OFF = ACC_BYTES;                     // slice records: 2 float4 per slice of a group
constexpr int MARK_OFF = REC_OFF + 16 * 8;             // 64 row marks
constexpr int TAG_OFF = MARK_OFF + 4 * 64;             // one byte per origin
constexpr int WAVE_LDS_BYTES = TAG_OFF + TAG_N;
#define SC_DLT 2e-3f   // margin on p: the FP32 position error is < 3e-4 for P <= 1024
#define SC_EPS 1e-3f   // margin on enumeration bounds
#define SC_RAD 6.95f   // 4 sqrt(3) + margin: |p - brick centre| for an origin in [lo, lo+7]^3

template <int A>
__device__ __forceinline__ int acc_slot(int i, int j, int k)
{
    return A == 0 ? (i * HALO + k) * HALO + j : (A == 1 ? (j * HALO + k) * HALO + i : (k * HALO + j) * HALO + i);
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave64 inclusive scans on DPP (GFX9 row shifts and row broadcasts): no
// LDS traffic.  Lanes without a source keep the identity.
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143
__device__ __forceinline__ int wave_scan_add(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(1), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(2), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(4), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(8), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST15, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST31, 0xc, 0xf, false);
    return v;
}
__device__ __forceinline__ int wave_scan_max(int v)
{
    const int lo = -2147483647 - 1;
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_SHR(1), 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_SHR(2), 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_SHR(4), 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_SHR(8), 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_BCAST15, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, DPP_ROW_BCAST31, 0xc, 0xf, false));
    return v;
}

// floor(sqrt(n)) for 0 <= n < 2^24
__device__ __forceinline__ int isqrt24(int n)
{
    int r = (int)__builtin_amdgcn_sqrtf((float)n);
    r -= (r * r > n);
    r += ((r + 1) * (r + 1) <= n);
    return r;
}

// gauss_2 / trilinear weights of the 8 corners (i0+di, j0+dj, k0+dk),
// corner index di + 2 dj + 4 dk; each equals corner_weight() bit for bit.
__device__ __forceinline__ void corner_weights8(int mode, int i0, int j0, int k0, float xx, float yy, float zz,
                                                float h, float weight, const float *etab, float wg[8])
{
    const float dx[2] = {(float)i0 - xx, (float)(i0 + 1) - xx};
    const float dy[2] = {(float)j0 - yy, (float)(j0 + 1) - yy};
    const float dz[2] = {(float)k0 - zz, (float)(k0 + 1) - zz};
    if (mode == 0) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const float sxy = dx[b & 1] * dx[b & 1] + dy[b >> 1] * dy[b >> 1];
#pragma unroll
            for (int dk = 0; dk < 2; dk++) {
                const float r = sxy + dz[dk] * dz[dk];       // Util::hypot3sq
                const float f = r * h;
                const int g = (int)__builtin_fma((double)f, 50.0, 0.5);  // Util::fast_exp index: f*50 is exact in double
                wg[b + 4 * dk] = etab[g] * weight;
            }
        }
    } else {
        double ax[2], ay[2], az[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            ax[t] = 1.0 - (double)fabsf(dx[t]);
            ay[t] = 1.0 - (double)fabsf(dy[t]);
            az[t] = 1.0 - (double)fabsf(dz[t]);
        }
#pragma unroll
        for (int cI = 0; cI < 8; cI++)
            wg[cI] = (float)(ax[cI & 1] * ay[(cI >> 1) & 1] * az[cI >> 2] * (double)weight);
    }
}

// A pixel waiting for its turn to add into the brick accumulator: value,
// position and weight (the corner weights are recomputed from these in the
// window that adds it).  Non-pending lanes hold a finite position, so the
// weight computation they run along with the wave stays inside the tables.
struct PendPix {
    float re, im, px, py, pz, pw;
    bool pend;
};

// Election of a window: one lane per stencil origin among the pending
// pixels wins the right to add its 8 corners; pixels carried over from
// earlier windows take precedence over the window's new ones (`fresh`), so a
// carried pixel is added in the next window it takes part in unless it
// collides with another carried pixel.  The losers stay pending.
__device__ __forceinline__ bool scatter_elect(const PendPix &st, bool fresh, const int lo[3], int lane,
                                              unsigned char *tag)
{
    const int i0 = (int)floorf(st.px) - lo[0], j0 = (int)floorf(st.py) - lo[1], k0 = (int)floorf(st.pz) - lo[2];
    const int ti = st.pend ? i0 + BRICK * (j0 + BRICK * k0) : 0;  // in [0, BRICK^3) while pending
    if (st.pend && fresh) tag[ti] = (unsigned char)lane;
    wave_sync_lds();
    if (st.pend && !fresh) tag[ti] = (unsigned char)lane;
    wave_sync_lds();
    const bool win = st.pend && tag[ti] == (unsigned char)lane;
    wave_sync_lds();
    return win;
}

// The winners add their 8 corners: lanes with different origins can still
// meet on a cell through different corners, so the corners go one after
// another with a wave-level ordering point between them.
template <int A>
__device__ __forceinline__ void scatter_add(const PendPix &st, bool win, const int lo[3], const InsertConst &c,
                                            float *acc, const float *etab)
{
    const int f0 = (int)floorf(st.px), f1 = (int)floorf(st.py), f2 = (int)floorf(st.pz);
    float wg[8];
    corner_weights8(c.mode, f0, f1, f2, st.px, st.py, st.pz, c.h, st.pw, etab, wg);
    const int cbase = acc_slot<A>(f0 - lo[0], f1 - lo[1], f2 - lo[2]);
    float *const p_re = acc + cbase, *const p_im = acc + ACC_N + cbase, *const p_w = acc + 2 * ACC_N + cbase;
#pragma unroll
    for (int cI = 0; cI < 8; cI++) {
        const int di = cI & 1, dj = (cI >> 1) & 1, dk = cI >> 2;
        if (win) {
            const int a = acc_slot<A>(di, dj, dk);
            p_re[a] += st.re * wg[cI];
            p_im[a] += st.im * wg[cI];
            p_w[a] += wg[cI];
        }
        wave_sync_lds();
    }
}

// Up to GROUP slices of an item in one stream of 64-pixel windows: lane
// (g, r) = (lane / 16, lane % 16) sets up row r of slice g (a stencil origin
// in [lo, lo+7]^3 puts p within 4 sqrt(3) of the brick centre, i.e. at most
// 14 rows), the rows' pixel counts are scanned over the wave, and each
// window gives the next pixels of the stream, whatever slice they belong to,
// to the lanes that hold no pending pixel.
constexpr int GROUP = 4;

// The slice of lane (g, r) in a group: index, t3d rows 0-1, weight.  Loaded
// one group ahead, so the reads are in flight while the previous group's
// windows run.
struct GroupSlice {
    int q;  // -1: no slice
    float mm[6];
    float wgt;
};

// Slice record of lane (g, r) for the group starting at item slice s: the
// index from the 64 list entries held in qbuf (entries qbase .. qbase+63).
__device__ __forceinline__ void fetch_group_slice(GroupSlice &gs, int s, int count, int qbuf, int qbase, int lane,
                                                  int nslices, const float *__restrict__ t3d,
                                                  const float *__restrict__ wq)
{
    const int idx = s + (lane >> 4);
    const int qi = __builtin_amdgcn_ds_bpermute((idx - qbase) << 2, qbuf);
    const int q = idx < count ? qi : -1;
    const bool qv = q >= 0 && q < nslices;  // never written list entries are skipped
    gs.q = qv ? q : -1;
#pragma unroll
    for (int d = 0; d < 6; d++) gs.mm[d] = qv ? t3d[9 * (size_t)q + d] : 0.f;
    gs.wgt = qv ? wq[q] : 0.f;
}

template <int A, bool SSNR>
__device__ __forceinline__ void scatter_group(const float2 *__restrict__ slices, int nsym, const GroupSlice &gs,
                                              const InsertConst &c, const int lo[3], int lane, float *acc,
                                              unsigned char *tag, int *mark, float4 *rec, const float *etab,
                                              const float *rtab, const float *__restrict__ rw, PendPix &st,
                                              bool &fresh)
{
    const int H = c.H, P = c.P;
    const unsigned slice_elems = (unsigned)P * (unsigned)(H + 1);
    const int g = lane >> 4, r = lane & 15;
    const int q = gs.q;
    const bool qv = q >= 0;
    float mm[6];
#pragma unroll
    for (int d = 0; d < 6; d++) mm[d] = gs.mm[d];
    const float wgt = gs.wgt;
    if (r == 0) {
        rec[2 * g] = make_float4(mm[0], mm[1], mm[2], mm[3]);
        rec[2 * g + 1] = make_float4(mm[4], mm[5], wgt, __int_as_float(qv ? q / nsym : 0));  // particle
    }
    // the pixel nearest the brick centre's projection; bounds are relative to it
    const float cx = (float)lo[0] + 4.0f, cy = (float)lo[1] + 4.0f, cz = (float)lo[2] + 4.0f;
    const float sc = cx * mm[0] + cy * mm[1] + cz * mm[2];
    const float tc = cx * mm[3] + cy * mm[4] + cz * mm[5];
    const int xc0 = (int)rintf(sc), yc0 = (int)rintf(tc);
    const int dya = (int)ceilf(tc - (float)yc0 - SC_RAD), dyb = (int)floorf(tc - (float)yc0 + SC_RAD);
    const int dxa = (int)ceilf(sc - (float)xc0 - SC_RAD), dxb = (int)floorf(sc - (float)xc0 + SC_RAD);
    // row dy: x range where p[d] in [lo, lo+8) (+- margin) on every axis
    int cnt = 0, xs = 0;
    const int y = yc0 + dya + r;
    if (qv && dya + r <= dyb && y >= -H && y <= H - 1) {
        const float dyf = (float)(dya + r);
        float xl = (float)dxa, xh = (float)dxb;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float ud = mm[d], vd = mm[3 + d];
            if (fabsf(ud) >= 0.05f) {
                const float P0 = (float)xc0 * ud + (float)yc0 * vd;
                const float iu = __builtin_amdgcn_rcpf(ud);
                const float t0 = ((float)lo[d] - P0 - SC_DLT - dyf * vd) * iu;
                const float t1 = ((float)(lo[d] + BRICK) - P0 + SC_DLT - dyf * vd) * iu;
                xl = fmaxf(xl, fminf(t0, t1));
                xh = fminf(xh, fmaxf(t0, t1));
            }
        }
        int a = max((int)ceilf(xl - SC_EPS), -xc0);
        int b = min((int)floorf(xh + SC_EPS), H - xc0);
        if (!c.corners) b = min(b, isqrt24(H * H + H - y * y) - xc0);  // hypot_fast_int(x,y) <= P/2
        xs = xc0 + a;
        cnt = max(0, b - a + 1);
    }
    const int incl = wave_scan_add(cnt);
    const int total = __builtin_amdgcn_readlane(incl, 63);
    const int excl = incl - cnt;
    // row record read by the slots of this row: (x - slot, y) as two int16
    const int rowrec = (int)(((unsigned)(xs - excl) & 0xffffu) | ((unsigned)y << 16));
    // Software-pipelined windows: the election of the pending pixels comes
    // first, then the lanes that will be free after this window load their
    // next pixels (slots next, next+1, ... in lane order), and the winners'
    // corner updates run while those slice reads are in flight.
    for (int next = 0; next < total;) {
        const bool win = scatter_elect(st, fresh, lo, lane, tag);
        const bool freed = !st.pend || win;
        const unsigned long long fm = __ballot(freed);
        const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(fm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)fm, 0u));
        // row of slot next+lane: the last non-empty row starting at or before it
        mark[lane] = -1;
        wave_sync_lds();
        if (cnt > 0 && excl >= next && excl < next + 64) mark[excl - next] = lane;
        if (cnt > 0 && excl < next && excl + cnt > next) mark[0] = lane;
        wave_sync_lds();
        const int rk = wave_scan_max(mark[lane]) & 63;
        const int rl = __builtin_amdgcn_ds_bpermute(rank << 2, rk);  // row of slot next+rank
        const int rr = __builtin_amdgcn_ds_bpermute(rl << 2, rowrec);
        const int s = next + rank;
        PendPix ns;
        ns.pend = false;
        bool ncj = false;
        if (freed && s < total) {
            const float4 ra = rec[2 * (rl >> 4)], rb = rec[2 * (rl >> 4) + 1];
            const int x = (int)(short)(rr & 0xffff) + s, ry = rr >> 16;
            const float pm[6] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y};
            float px, py, pz;
            pixel_coords(x, ry, pm, rtab, H, c.Pf, px, py, pz);
            const int f0 = (int)floorf(px), f1 = (int)floorf(py), f2 = (int)floorf(pz);
            // x in [0, P/2] holds by construction; origins lie in [-P/2-1, ...): the
            // ones past P/2 only reach cells that are never merged
            if ((unsigned)x <= (unsigned)H && (unsigned)(f0 - lo[0]) < (unsigned)BRICK &&
                (unsigned)(f1 - lo[1]) < (unsigned)BRICK && (unsigned)(f2 - lo[2]) < (unsigned)BRICK) {
                // slice_value() with the conjugation applied after the corner
                // updates, so that nothing waits for the read before then
                const bool cj = x == 0 && ry < 0;
                const int row = cj ? -ry : (ry < 0 ? P + ry : ry);
                const float2 dt = (slices + (unsigned long long)(unsigned)__float_as_int(rb.w) *
                                                slice_elems)[row * (H + 1) + x];
                ns.re = dt.x;
                ns.im = dt.y;
                ncj = cj;
                ns.px = px;
                ns.py = py;
                ns.pz = pz;
                // class_ssnr weighting: the particle's weight at radius hypot_fast_int(x, y)
                ns.pw = SSNR ? rw[(unsigned long long)(unsigned)__float_as_int(rb.w) * (unsigned)(H + 1) +
                                  (unsigned)hypot_int(x, ry)]
                             : rb.z;
                ns.pend = true;
            }
        }
        wave_sync_lds();  // rec reads before the next group rewrites them
        if (__ballot(win)) scatter_add<A>(st, win, lo, c, acc, etab);
        if (freed) {
            st.pend = ns.pend;
            if (ns.pend) {
                st = ns;
                if (ncj) st.im = -st.im;
            }
        }
        fresh = freed && ns.pend;
        next += __popcll(fm);
    }
}

// Partial brick layout in HBM: [field re/im/w][z][y][x] of the brick's 9^3 cells.
template <int A, bool SSNR>
__device__ __forceinline__ void scatter_item(const float2 *__restrict__ slices, int nslices, int nsym,
                                             const float *__restrict__ t3d, const float *__restrict__ wq,
                                             const InsertConst &c, const int *__restrict__ list, const WorkItem &w,
                                             const int lo[3], int lane, float *acc, unsigned char *tag, int *mark,
                                             float4 *rec, const float *etab, const float *rtab,
                                             const float *__restrict__ rw, float *out)
{
    PendPix st;
    st.re = st.im = st.pw = 0.f;
    st.px = (float)lo[0];
    st.py = (float)lo[1];
    st.pz = (float)lo[2];
    st.pend = false;
    bool fresh = false;
    // list entries 64 at a time; each group's slice records fetched one group ahead
    int qbase = 0;
    int qbuf = lane < w.count ? list[w.start + lane] : -1;
    GroupSlice nxt;
    fetch_group_slice(nxt, 0, w.count, qbuf, qbase, lane, nslices, t3d, wq);
    for (int s0 = 0; s0 < w.count; s0 += GROUP) {
        const GroupSlice cur = nxt;
        const int s1 = s0 + GROUP;
        if (s1 < w.count) {
            if (s1 + GROUP > qbase + 64) {
                qbase = s1;
                qbuf = qbase + lane < w.count ? list[w.start + qbase + lane] : -1;
            }
            fetch_group_slice(nxt, s1, w.count, qbuf, qbase, lane, nslices, t3d, wq);
        }
        scatter_group<A, SSNR>(slices, nsym, cur, c, lo, lane, acc, tag, mark, rec, etab, rtab, rw, st, fresh);
    }
    while (__ballot(st.pend)) {
        const bool win = scatter_elect(st, fresh, lo, lane, tag);
        scatter_add<A>(st, win, lo, c, acc, etab);
        st.pend = st.pend && !win;
        fresh = false;
    }
    wave_sync_lds();
#pragma unroll
    for (int wv = 0; wv < (ACC_N + 63) / 64; wv++) {
        const int cc = lane + 64 * wv;
        if (cc < ACC_N) {
            const int z = cc / (HALO * HALO), yx = cc - z * HALO * HALO, y = yx / HALO, x = yx - y * HALO;
            const int a = acc_slot<A>(x, y, z);
            out[cc] = acc[a];
            out[ACC_N + cc] = acc[ACC_N + a];
            out[2 * ACC_N + cc] = acc[2 * ACC_N + a];
        }
    }
    wave_sync_lds();
}

// SSNR: per-pixel weights from the particles' radial class_ssnr tables rw
// ((P/2+1) floats per particle) instead of the slice weight.
template <bool SSNR>
__global__ void __launch_bounds__(256) k_scatter_insert(const float2 *__restrict__ slices, int nslices, int nsym,
                                                        const float *__restrict__ t3d, const float *__restrict__ wq,
                                                        InsertConst c, BrickGeom g, BinBuffers bb,
                                                        const float *__restrict__ exptab_g,
                                                        const float *__restrict__ rtab_g,
                                                        const float *__restrict__ rw)
{
    extern __shared__ float4 lds4[];
    char *base = reinterpret_cast<char *>(lds4);
    const int lane = threadIdx.x & 63;
    char *mine = base + (size_t)(threadIdx.x >> 6) * WAVE_LDS_BYTES;
    float *acc = reinterpret_cast<float *>(mine);
    float4 *rec = reinterpret_cast<float4 *>(mine + REC_OFF);
    int *mark = reinterpret_cast<int *>(mine + MARK_OFF);
    unsigned char *tag = reinterpret_cast<unsigned char *>(mine + TAG_OFF);
    float *etab = reinterpret_cast<float *>(base + (size_t)(blockDim.x >> 6) * WAVE_LDS_BYTES);
    float *rtab = etab + EXPN;  // P+2 entries: fl(i/P), i in [-H, H+1]
    for (int i = threadIdx.x; i < EXPN; i += blockDim.x) etab[i] = exptab_g[i];
    for (int i = threadIdx.x; i < c.P + 2; i += blockDim.x) rtab[i] = rtab_g[i];
    __syncthreads();
    const int H = c.H, nb = g.nb;
    const int total = bb.ctl[1];
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(&bb.ctl[0], 1);
        it = __shfl(it, 0, 64);
        if (it >= total) break;
        const WorkItem w = bb.items[it];
        const int brick = w.key / 3, A = w.key - 3 * brick;
        const int lo[3] = {-H - 1 + BRICK * (brick % nb), -H - 1 + BRICK * ((brick / nb) % nb),
                           -H - 1 + BRICK * (brick / (nb * nb))};
#pragma unroll
        for (int wv = 0; wv < (3 * ACC_N + 63) / 64; wv++)
            if (lane + 64 * wv < 3 * ACC_N) acc[lane + 64 * wv] = 0.f;
        wave_sync_lds();
        float *out = bb.part + (size_t)it * PART_FLOATS;
        if (A == 0) scatter_item<0, SSNR>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, tag, mark, rec, etab, rtab, rw, out);
        else if (A == 1) scatter_item<1, SSNR>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, tag, mark, rec, etab, rtab, rw, out);
        else scatter_item<2, SSNR>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, tag, mark, rec, etab, rtab, rw, out);
    }
}

size_t scatter_lds_bytes(int P, int waves_per_block)
{
    return (size_t)waves_per_block * WAVE_LDS_BYTES + sizeof(float) * (EXPN + P + 2);
}

hipError_t launch_scatter_insert(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                                 InsertConst c, BrickGeom g, BinBuffers b, const float *exptab, const float *rtab,
                                 const float *rw, int grid, int waves_per_block, hipStream_t s)
{
    if (waves_per_block < 1 || waves_per_block > 4) return hipErrorInvalidValue;
    if (rw)
        hipLaunchKernelGGL(k_scatter_insert<true>, dim3(grid), dim3(64 * waves_per_block),
                           scatter_lds_bytes(c.P, waves_per_block), s, slices, nslices, nsym, t3d, wq, c, g, b, exptab,
                           rtab, rw);
    else
        hipLaunchKernelGGL(k_scatter_insert<false>, dim3(grid), dim3(64 * waves_per_block),
                           scatter_lds_bytes(c.P, waves_per_block), s, slices, nslices, nsym, t3d, wq, c, g, b, exptab,
                           rtab, rw);
    return hipGetLastError();
}

// ------------------------------------------------------------ wide-stencil brick scatter
// The inserters with stencils of S = 3 to 9 cells per axis (gauss_var,
// gauss_5, gauss_3, kaiser_bessel, kaiser_bessel_derived, gridding_5,
// gridding_7 and fourier_iter's Gaussian) on bricks of 4^3 stencil origins,
// binned like gauss_2's on the origin grid of the mode.  A work item's wave
// owns the brick's (3 + S)^3 cells in LDS.  For each slice of the item it
// tests the 8 x 8 pixels around the projection of the brick's centre (a
// pixel whose origin lies in the brick is within 2 sqrt 3 of that centre in
// p, so within 7 pixels per axis in the slice), then adds the stencils of the
// pixels whose origin lies in the brick one pixel after another, all 64 lanes
// over the pixel's S^3 cells.  The cells of one stencil are distinct and the
// pixels of a wave follow one another in program order, so plain LDS
// read-modify-writes suffice: no atomics and no election.  Each contribution
// is computed by wide_setup / wide_term, as in every other form.
#define WR_CAND 3.4691f  // 2 sqrt 3 + margin for the FP32 position and origin rounding

__device__ __forceinline__ float bcast_f(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

template <int MODE>
__global__ void __launch_bounds__(256) k_scatter_wide(const float2 *__restrict__ slices, int nslices, int nsym,
                                                      const float *__restrict__ t3d, const float *__restrict__ wq,
                                                      InsertConst c, BrickGeom g, BinBuffers bb,
                                                      const float *__restrict__ exptab_g,
                                                      const float *__restrict__ rtab, const float *__restrict__ rw)
{
    constexpr int S = wide_cells(MODE), HL = WBS + S - 1, HC = HL * HL * HL;
    constexpr int NC = S * S * S, NIT = (NC + 63) / 64;
    // the Gaussian 6^3 stencils (gauss_var, gauss_5, fourier_iter without a reference): lanes 0..53
    // keep one x offset oi = lane % 6 in every pass and take the (oj, ok) pair lane / 6 + 9 t in
    // pass t, so the x distance is computed once per pixel
    constexpr bool G6 = MODE == 2 || MODE == 3 || MODE == 10;
    extern __shared__ float lds_w[];
    float *etab = lds_w;  // the 1000 fast_exp entries
    float *acc = lds_w + 1024 + (size_t)(threadIdx.x >> 6) * 3 * HC;  // re, im, w over [z][y][x] of the cells
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < 1000; i += blockDim.x) etab[i] = exptab_g[i];
    __syncthreads();
    const int H = c.H, P = c.P, nb = g.nb;
    // this lane's stencil cell in pass t: offsets oi | oj << 4 | ok << 8 (-1 past S^3) and its LDS offset
    int oc[NIT], off[NIT];
    float fo_j[NIT], fo_k[NIT];
    const float fo_i = (float)(lane % S);
#pragma unroll
    for (int t = 0; t < NIT; t++) {
        int oi, oj, ok;
        bool in;
        if (G6) {
            const int pr = lane / 6 + 9 * t;
            oi = lane % 6;
            oj = pr % 6;
            ok = pr / 6;
            in = lane < 54;
        } else {
            const int ci = lane + 64 * t;
            oi = ci % S;
            oj = (ci / S) % S;
            ok = ci / (S * S);
            in = ci < NC;
        }
        oc[t] = in ? (oi | (oj << 4) | (ok << 8)) : -1;
        off[t] = in ? (ok * HL + oj) * HL + oi : 0;
        fo_j[t] = (float)oj;
        fo_k[t] = (float)ok;
    }
    const int total = bb.ctl[1];
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(&bb.ctl[0], 1);
        it = __shfl(it, 0, 64);
        if (it >= total) break;
        const WorkItem w = bb.items[it];
        const int brick = w.key / 3;
        const int bc[3] = {brick % nb, (brick / nb) % nb, brick / (nb * nb)};
        int lo[3];
        float pc[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lo[d] = g.ow + WBS * bc[d];
            pc[d] = g.po + (float)(WBS * bc[d]) + 0.5f * WBS;  // exact: half-integers of magnitude < 2^20
        }
        for (int i = lane; i < 3 * HC; i += 64) acc[i] = 0.f;
        wave_sync_lds();
        int qbuf = -1;
        for (int s = 0; s < w.count; s++) {
            if ((s & 63) == 0) qbuf = s + lane < w.count ? bb.list[w.start + s + lane] : -1;
            const int q = __builtin_amdgcn_readlane(qbuf, s & 63);
            if (q < 0 || q >= nslices) continue;  // never written list entries are skipped
            const float wgt = wq[q];
            if (wgt == 0.0f) continue;
            float mm[6];
#pragma unroll
            for (int d = 0; d < 6; d++) mm[d] = t3d[9 * (size_t)q + d];
            const float sc = pc[0] * mm[0] + pc[1] * mm[1] + pc[2] * mm[2];
            const float tc = pc[0] * mm[3] + pc[1] * mm[4] + pc[2] * mm[5];
            const int x = (int)ceilf(sc - WR_CAND) + (lane & 7), y = (int)ceilf(tc - WR_CAND) + (lane >> 3);
            bool ok = x >= 0 && x <= H && y >= -H && y <= H - 1 &&
                      (c.corners || x * x + y * y <= c.rcut * c.rcut + c.rcut);  // hypot_fast_int(x, y) <= rcut
            WidePix px;
            int pk = 0;
            if (ok) {
                const int p = q / nsym;
                const float weight = rw ? rw[(size_t)p * (H + 1) + hypot_int(x, y)] : wgt;
                float xx, yy, zz;
                pixel_coords(x, y, mm, rtab, H, c.Pf, xx, yy, zz);
                ok = wide_setup(MODE, xx, yy, zz, weight, H, etab, px, c.ref);
                if (ok) {
#pragma unroll
                    for (int d = 0; d < 3; d++) ok = ok && (unsigned)(px.a[d] - lo[d]) < (unsigned)WBS;
                }
                if (ok) {
                    const float2 dt = slice_value(slices + (size_t)p * P * (H + 1), x, y, P, H);
                    px.dre = dt.x;
                    px.dim = dt.y;
#pragma unroll
                    for (int d = 0; d < 3; d++)
                        pk |= ((px.a[d] - lo[d]) | ((px.lo[d] - px.a[d]) << 2) | ((px.hi[d] - px.a[d]) << 6))
                              << (10 * d);
                }
            }
            unsigned long long hits = __ballot(ok);
            while (hits) {
                const int j = (int)__builtin_ctzll(hits);
                hits &= hits - 1;
                WidePix u;
                u.xx = bcast_f(px.xx, j);
                u.yy = bcast_f(px.yy, j);
                u.zz = bcast_f(px.zz, j);
                u.w = bcast_f(px.w, j);
                u.ww = bcast_f(px.ww, j);
                u.h = bcast_f(px.h, j);
                u.dre = bcast_f(px.dre, j);
                u.dim = bcast_f(px.dim, j);
                u.nre = MODE == 11 ? bcast_f(px.nre, j) : 0.f;
                u.nim = MODE == 11 ? bcast_f(px.nim, j) : 0.f;
                u.H = H;
                u.ref = c.ref;
                const int pkj = __builtin_amdgcn_readlane(pk, j);
                bool full = true;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int f = pkj >> (10 * d);
                    u.a[d] = lo[d] + (f & 3);
                    u.lo[d] = u.a[d] + ((f >> 2) & 15);
                    u.hi[d] = u.a[d] + ((f >> 6) & 15);
                    full = full && u.lo[d] == u.a[d] && u.hi[d] == u.a[d] + S - 1;
                }
                const int base = ((u.a[2] - lo[2]) * HL + (u.a[1] - lo[1])) * HL + (u.a[0] - lo[0]);
                // the pixel's cells are distinct: every lane reads its cells first (lanes past the
                // stencil read cell `base`, in bounds) and the lanes with a contribution write them
                // back after the terms are computed
                float ar[NIT], ai[NIT], aw[NIT];
#pragma unroll
                for (int t = 0; t < NIT; t++) {
                    ar[t] = acc[base + off[t]];
                    ai[t] = acc[HC + base + off[t]];
                    aw[t] = acc[2 * HC + base + off[t]];
                }
                float vr[NIT], vi[NIT], vg[NIT];
                bool on[NIT];
#pragma unroll
                for (int t = 0; t < NIT; t++) {
                    const int o = oc[t];
                    const int i = u.a[0] + (o & 15), jj = u.a[1] + ((o >> 4) & 15), k = u.a[2] + ((o >> 8) & 15);
                    on[t] = o >= 0 && (full || (i >= u.lo[0] && i <= u.hi[0] && jj >= u.lo[1] && jj <= u.hi[1] &&
                                                k >= u.lo[2] && k <= u.hi[2]));
                    if (!G6 && on[t]) on[t] = wide_term(MODE, u, i, jj, k, etab, vr[t], vi[t], vg[t]);
                }
                if (G6) {
                    // wide_term's Gaussian, pass by pass: (float)i = (float)a + (float)o exactly (small
                    // integers), r = hypot3sq in the reference's order, fast_exp(-r h): the table entry
                    // (int)(-(double)f * 50 + 0.5) -- f * 50 is exact in double, so one fma -- for
                    // f >= -19.98, expf below (f = -(r h) <= 0 since h > 0).  All passes' table reads go
                    // out before the first is used; expf runs only when a lane of the wave needs it.
                    const float dx = ((float)u.a[0] + fo_i) - u.xx;
                    const float dx2 = dx * dx;
                    float fv[NIT], tv[NIT];
                    bool big[NIT], anyb = false;
#pragma unroll
                    for (int t = 0; t < NIT; t++) {
                        const float dy = ((float)u.a[1] + fo_j[t]) - u.yy, dz = ((float)u.a[2] + fo_k[t]) - u.zz;
                        const float f = -(dx2 + dy * dy + dz * dz) * u.h;
                        fv[t] = f;
                        big[t] = f < -19.98f;  // the same floats as (double)f < -19.98
                        anyb = anyb || big[t];
                        const int gi = big[t] ? 0 : (int)__builtin_fma(-(double)f, 50.0, 0.5);
                        tv[t] = etab[gi];
                    }
                    if (__ballot(anyb)) {
#pragma unroll
                        for (int t = 0; t < NIT; t++)
                            if (big[t]) tv[t] = expf(fv[t]);
                    }
#pragma unroll
                    for (int t = 0; t < NIT; t++) {
                        const float gg = tv[t];
                        vg[t] = gg * u.ww;
                        vr[t] = u.dre * gg * u.ww;
                        vi[t] = u.dim * gg * u.ww;
                        on[t] = on[t] && !(fabsf(vg[t]) < FLT_MIN);
                    }
                }
#pragma unroll
                for (int t = 0; t < NIT; t++)
                    if (on[t]) {
                        acc[base + off[t]] = ar[t] + vr[t];
                        acc[HC + base + off[t]] = ai[t] + vi[t];
                        acc[2 * HC + base + off[t]] = aw[t] + vg[t];
                    }
            }
        }
        wave_sync_lds();
        float *out = bb.part + (size_t)it * g.part_floats;
        for (int i = lane; i < 3 * HC; i += 64) out[i] = acc[i];
        wave_sync_lds();
    }
}

size_t scatter_wide_lds_bytes(int mode, int waves_per_block)
{
    const int HL = WBS + wide_cells(mode) - 1;
    return sizeof(float) * (1024 + (size_t)waves_per_block * 3 * HL * HL * HL);
}

hipError_t launch_scatter_wide(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                               InsertConst c, BrickGeom g, BinBuffers b, const float *exptab, const float *rtab,
                               const float *rw, int grid, int waves_per_block, hipStream_t s)
{
    const int S = wide_cells(c.mode);
    if (S == 0 || waves_per_block < 1 || waves_per_block > 4 || g.bs != WBS || g.halo != WBS + S - 1)
        return hipErrorInvalidValue;
    const size_t lds = scatter_wide_lds_bytes(c.mode, waves_per_block);
    if (lds > 65536) return hipErrorInvalidValue;
    const dim3 gr(grid), bl(64 * waves_per_block);
#define FH_WIDE(M) \
    case M: hipLaunchKernelGGL(k_scatter_wide<M>, gr, bl, lds, s, slices, nslices, nsym, t3d, wq, c, g, b, exptab, rtab, rw); break;
    switch (c.mode) {
        FH_WIDE(2)
        FH_WIDE(3)
        FH_WIDE(5)
        FH_WIDE(6)
        FH_WIDE(7)
        FH_WIDE(8)
        FH_WIDE(9)
        FH_WIDE(10)
        FH_WIDE(11)
    default:
        return hipErrorInvalidValue;
    }
#undef FH_WIDE
    return hipGetLastError();
}

// ------------------------------------------------------------ merge of partial bricks
// Sum over the work items of the bricks whose cells include unfolded cell
// (i, j, k): brick b covers cells [ow + BS*b, ow + BS*b + HL) per axis, so a
// cell lies in the bricks floor(u / BS) down to ceil((u - HL + 1) / BS),
// u = cell - ow (for gauss_2, BS 8 and HL 9: the brick owning the origin and,
// on each axis where the cell is a brick's first, also the previous brick).
template <int BS, int HL>
__device__ __forceinline__ void partial_sum(const float *__restrict__ part, const int *__restrict__ item_off, int nb,
                                            int ow, int pf, int i, int j, int k, float &re, float &im, float &w)
{
    constexpr int HC = HL * HL * HL;
    const int ui = i - ow, uj = j - ow, uk = k - ow;  // >= 0 for cells >= -H
    const int bi1 = min(ui / BS, nb - 1), bj1 = min(uj / BS, nb - 1), bk1 = min(uk / BS, nb - 1);
    const int bi0 = ui < HL ? 0 : (ui - HL + BS) / BS, bj0 = uj < HL ? 0 : (uj - HL + BS) / BS;
    const int bk0 = uk < HL ? 0 : (uk - HL + BS) / BS;
    re = im = w = 0.f;
    for (int bk = bk1; bk >= bk0; bk--)
        for (int bj = bj1; bj >= bj0; bj--)
            for (int bi = bi1; bi >= bi0; bi--) {
                const int b = (bk * nb + bj) * nb + bi;
                const int cc = ((uk - BS * bk) * HL + (uj - BS * bj)) * HL + (ui - BS * bi);
                const int i0 = item_off[3 * b], i1 = item_off[3 * b + 3];
                for (int it = i0; it < i1; it++) {
                    const float *p = part + (size_t)it * pf + cc;
                    re += p[0];
                    im += p[HC];
                    w += p[2 * HC];
                }
            }
}

// Storage voxel (xs, ys, zs) of the half volume receives, per
// add_complex_at_fast (libEM/emdata_core.h:368-413): the direct cells
// (xs, j, k) with wrap(j) = ys, wrap(k) = zs; the conjugate cells (-xs, j, k)
// with wrapneg(j) = ys, wrapneg(k) = zs; on the x = 0 and x = P/2 planes the
// conjugate copies of the double insertion (value only) and, at x = P/2,
// the x = -P/2 cells added twice with their weight once; at (0,0,0) and
// (P/2,0,0) the real part and weight with the imaginary part set to 0.
template <int BS, int HL>
__global__ void __launch_bounds__(256) k_merge_partials(const float *__restrict__ part, const int *__restrict__ item_off,
                                                        BrickGeom g, InsertConst c, float2 *data, float *norm)
{
    const int H = c.H, P = c.P, nb = g.nb, ow = g.ow, pf = g.part_floats;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)P * P * (H + 1)) return;
    const int xs = (int)(idx % (H + 1));
    const long long rest = idx / (H + 1);
    const int ys = (int)(rest % P), zs = (int)(rest / P);
    float r, i, w;
    if (ys == 0 && zs == 0 && (xs == 0 || xs == H)) {
        partial_sum<BS, HL>(part, item_off, nb, ow, pf, xs, 0, 0, r, i, w);
        if (xs == H) {
            float r2, i2, w2;
            partial_sum<BS, HL>(part, item_off, nb, ow, pf, -H, 0, 0, r2, i2, w2);
            r += r2;
            i += i2;
            w += w2;
        }
        if (r == 0.f && i == 0.f && w == 0.f) return;
        float2 d = data[idx];
        d.x += r;
        d.y = 0.f;
        data[idx] = d;
        norm[idx] += w;
        return;
    }
    int jd[2], njd = 0, jn[2], njn = 0, kd[2], nkd = 0, kn[2], nkn = 0;
    if (ys <= H) jd[njd++] = ys;
    if (ys >= H) jd[njd++] = ys - P;
    if (ys <= H) jn[njn++] = -ys;
    if (ys >= H && P - ys <= H && P - ys != -ys) jn[njn++] = P - ys;
    if (zs <= H) kd[nkd++] = zs;
    if (zs >= H) kd[nkd++] = zs - P;
    if (zs <= H) kn[nkn++] = -zs;
    if (zs >= H && P - zs <= H && P - zs != -zs) kn[nkn++] = P - zs;
    float dre = 0.f, dim = 0.f, dw = 0.f;
    for (int a = 0; a < njd; a++)
        for (int b = 0; b < nkd; b++) {
            partial_sum<BS, HL>(part, item_off, nb, ow, pf, xs, jd[a], kd[b], r, i, w);
            dre += r;
            dim += i;
            dw += w;
        }
    for (int a = 0; a < njn; a++)
        for (int b = 0; b < nkn; b++) {
            if (xs == 0 || xs == H) {
                partial_sum<BS, HL>(part, item_off, nb, ow, pf, xs, jn[a], kn[b], r, i, w);
                dre += r;
                dim -= i;
                if (xs == H) {
                    partial_sum<BS, HL>(part, item_off, nb, ow, pf, -H, jn[a], kn[b], r, i, w);
                    dre += 2.f * r;
                    dim -= 2.f * i;
                    dw += w;
                }
            } else {
                partial_sum<BS, HL>(part, item_off, nb, ow, pf, -xs, jn[a], kn[b], r, i, w);
                dre += r;
                dim -= i;
                dw += w;
            }
        }
    if (dre == 0.f && dim == 0.f && dw == 0.f) return;
    float2 d = data[idx];
    d.x += dre;
    d.y += dim;
    data[idx] = d;
    norm[idx] += dw;
}

hipError_t launch_merge_partials(BrickGeom g, BinBuffers b, InsertConst c, float2 *data, float *norm, hipStream_t s)
{
    const long long n = (long long)c.P * c.P * (c.H + 1);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (g.bs == BRICK && g.halo == HALO)
        hipLaunchKernelGGL((k_merge_partials<BRICK, HALO>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else if (g.bs == WBS && g.halo == WBS + 2)
        hipLaunchKernelGGL((k_merge_partials<WBS, WBS + 2>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else if (g.bs == WBS && g.halo == WBS + 4)
        hipLaunchKernelGGL((k_merge_partials<WBS, WBS + 4>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else if (g.bs == WBS && g.halo == WBS + 5)
        hipLaunchKernelGGL((k_merge_partials<WBS, WBS + 5>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else if (g.bs == WBS && g.halo == WBS + 6)
        hipLaunchKernelGGL((k_merge_partials<WBS, WBS + 6>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else if (g.bs == WBS && g.halo == WBS + 8)
        hipLaunchKernelGGL((k_merge_partials<WBS, WBS + 8>), grid, dim3(256), 0, s, b.part, b.item_off, g, c, data, norm);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh

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
//  k_insert_plain  one launch per (slice, symmetry op), one thread per pixel,
//                  global atomics per contribution (the plain form).
//  k_bin_*         bins (slice, brick) pairs over the unfolded frequency grid
//                  [-P/2, P/2]^3 cut into BS^3 bricks.
//  k_brick_insert  persistent brick-stationary kernel: each work item holds a
//                  brick's accumulator in LDS, pulls the slice pixels whose
//                  2x2x2 stencil reaches the brick, and flushes once with the
//                  Hermitian fold of add_complex_at_fast as global atomics.
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
    if (x == 0 && y < 0) return conjf2(S[(size_t)(-y) * (H + 1)]);
    const int row = y < 0 ? P + y : y;
    return S[(size_t)row * (H + 1) + x];
}

// gauss_2 / trilinear weight of corner (i,j,k) for a pixel at (xx,yy,zz).
__device__ __forceinline__ float corner_weight(int mode, int i, int j, int k, float xx, float yy, float zz, float h,
                                               float weight, const float *etab)
{
    const float dx = (float)i - xx, dy = (float)j - yy, dz = (float)k - zz;
    if (mode == 0) {
        const float r = dx * dx + dy * dy + dz * dz;  // Util::hypot3sq
        const float f = r * h;                          // -(-r*h)
        const int g = (int)((double)f * 50.0 + 0.5);    // Util::fast_exp index
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

// ------------------------------------------------------------ plain form
__global__ void __launch_bounds__(256) k_insert_plain(const float2 *__restrict__ S, InsertConst c, Mat9 m,
                                                      float weight, float2 *data, float *norm,
                                                      const float *__restrict__ exptab,
                                                      const float *__restrict__ rtab_g)
{
    const int H = c.H, P = c.P;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = (int)blockIdx.y - H;
    if (x > H) return;
    if (!c.corners && x * x + y * y > H * H + H) return;  // hypot_fast_int(x,y) > P/2
    float xx, yy, zz;
    pixel_coords(x, y, m.m, rtab_g, H, c.Pf, xx, yy, zz);
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

hipError_t launch_insert_plain(const float2 *slice, InsertConst c, const float *t3d9, float weight, float2 *data,
                               float *norm, const float *exptab, const float *rtab, hipStream_t s)
{
    Mat9 m;
    for (int i = 0; i < 9; i++) m.m[i] = t3d9[i];
    dim3 grid((c.H + 1 + 63) / 64, c.P);
    hipLaunchKernelGGL(k_insert_plain, grid, dim3(64), 0, s, slice, c, m, weight, data, norm, exptab, rtab);
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
                    const float r = __fsqrt_rn((float)(x * x + y * y));
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

// One block row per (particle, symmetry op) slice; eight double sums per
// particle accumulated with float64 atomics into acc[p][8].
__global__ void __launch_bounds__(256) k_agree_batch(const float2 *__restrict__ slices, int nsym,
                                                     const float *__restrict__ t3d, const float *__restrict__ wq,
                                                     InsertConst c, float rmax, int sub,
                                                     const float2 *__restrict__ data, const float *__restrict__ norm,
                                                     const float *__restrict__ exptab, const float *__restrict__ rtab,
                                                     double *__restrict__ acc_out)
{
    const int H = c.H, P = c.P;
    const int q = blockIdx.y, p = q / nsym;
    const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = H + 2;
    const float2 *S = slices + (size_t)p * P * (H + 1);
    const float *t3dp = t3d + 9 * (size_t)p * nsym;
    const float weight = wq[q];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
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
                const float *Sf = reinterpret_cast<const float *>(S);
                const int slen = (P + 2) * P;
                const int fidx = 2 * x + (P + 2) * (y < 0 ? P + y : y);
                const float d20 = fidx < slen ? Sf[fidx] : 0.f;
                const float d21 = fidx + 1 < slen ? Sf[fidx + 1] : 0.f;
                const int x0 = round_ref(xx), y0 = round_ref(yy), z0 = round_ref(zz);
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
                    own_contribution(S, t3dp, nsym, xs, ysn, zsn, c, weight, exptab, rtab, ore, oim, ow);
                    if (inb) {
                        vr = vr - ore;
                        vi = vi - oim;
                    }
                    nv = nv - ow;
                }
                if (x0 < 0) vi = -vi;  // get_complex_at conjugate for x<0
                if (nv != 0.f) {
                    const float dt0 = __fdiv_rn(vr, nv), dt1 = __fdiv_rn(vi, nv), dt2 = nv;
                    const float r = __fsqrt_rn((float)(x * x + y * y));
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
    if (threadIdx.x < 8) unsafeAtomicAdd(&acc_out[(size_t)p * 8 + threadIdx.x], red[threadIdx.x][0]);
}

hipError_t launch_agree_batch(const float2 *slices, int nsym, const float *t3d, const float *wq, int nslices,
                              InsertConst c, float rmax, int sub, const float2 *data, const float *norm,
                              const float *exptab, const float *rtab, double *acc, hipStream_t s)
{
    if (nslices <= 0) return hipSuccess;
    c.corners = 0;
    hipLaunchKernelGGL(k_agree_batch, dim3(agree_blocks(c.P), nslices), dim3(256), 0, s, slices, nsym, t3d, wq, c, rmax,
                       sub, data, norm, exptab, rtab, acc);
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
// Brick b = (bx, by, bz) covers unfolded voxels [-H + BS*b, -H + BS*b + BS);
// a pixel reaches it when its stencil origin floor(p) lies in
// [lo-1, lo+BS-1], i.e. p in [lo-1, lo+BS) on every axis.  Pixel positions
// lie on the plane n.p = 0 (n = u x v of the slice's t3d rows), so per brick
// column along the dominant normal axis at most 5 bricks qualify.
struct SliceFrame {
    double u[3], v[3], n[3];
    int a, b1, b2;
};

__device__ __forceinline__ void slice_frame(const float *m, SliceFrame &f)
{
    for (int d = 0; d < 3; d++) {
        f.u[d] = m[d];
        f.v[d] = m[3 + d];
    }
    f.n[0] = f.u[1] * f.v[2] - f.u[2] * f.v[1];
    f.n[1] = f.u[2] * f.v[0] - f.u[0] * f.v[2];
    f.n[2] = f.u[0] * f.v[1] - f.u[1] * f.v[0];
    const double a0 = fabs(f.n[0]), a1 = fabs(f.n[1]), a2 = fabs(f.n[2]);
    f.a = (a0 >= a1 && a0 >= a2) ? 0 : (a1 >= a2 ? 1 : 2);
    f.b1 = f.a == 0 ? 1 : 0;
    f.b2 = f.a == 2 ? 1 : 2;
}

// Does the slice's pixel domain (half-disk x>=0 of radius sqrt(H^2+H), or
// the half-square with corners) come within reach of brick region R?
__device__ __forceinline__ bool brick_in_domain(const SliceFrame &f, const int bc[3], int H, int BS, int corners)
{
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = -H + BS * bc[d] - 1 + 0.5 * (BS + 1);
    const double s = c[0] * f.u[0] + c[1] * f.u[1] + c[2] * f.u[2];
    const double t = c[0] * f.v[0] + c[1] * f.v[1] + c[2] * f.v[2];
    const double rh = 0.8660254037844386 * (BS + 1) + 0.01;
    if (s < -rh) return false;
    if (corners) return s <= H + rh && t >= -H - rh && t <= H - 1 + rh;
    const double rr = sqrt((double)H * H + H) + rh;
    return s * s + t * t <= rr * rr;
}

template <int MODE>  // 0 count, 1 fill
__global__ void __launch_bounds__(256) k_bin(const float *__restrict__ t3d, const float *__restrict__ wq, int nslices,
                                             InsertConst c, BrickGeom g, BinBuffers bb)
{
    const int nb = g.nb, BS = g.Bs, H = c.H;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = blockIdx.y;
    if (col >= nb * nb || q >= nslices) return;
    if (wq[q] == 0.0f) return;  // weight 0 slices are skipped (reconstructor.cpp:1088)
    SliceFrame f;
    slice_frame(t3d + 9 * (size_t)q, f);
    const int c1 = col % nb, c2 = col / nb;
    const double dlt = 1e-3;
    const double L1 = -H + BS * c1 - 1 - dlt, U1 = -H + BS * c1 + BS + dlt;
    const double L2 = -H + BS * c2 - 1 - dlt, U2 = -H + BS * c2 + BS + dlt;
    const double na = f.n[f.a], n1 = f.n[f.b1], n2 = f.n[f.b2];
    double amin = 1e30, amax = -1e30;
    const double p1s[2] = {L1, U1}, p2s[2] = {L2, U2};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            const double pa = -(n1 * p1s[i] + n2 * p2s[j]) / na;
            amin = fmin(amin, pa);
            amax = fmax(amax, pa);
        }
    int lo = (int)floor((amin - dlt + H - BS) / BS) + 1;
    int hi = (int)floor((amax + dlt + H + 1) / BS);
    if (lo < 0) lo = 0;
    if (hi > nb - 1) hi = nb - 1;
    for (int ca = lo; ca <= hi; ca++) {
        int bc[3];
        bc[f.a] = ca;
        bc[f.b1] = c1;
        bc[f.b2] = c2;
        if (!brick_in_domain(f, bc, H, BS, c.corners)) continue;
        const int brick = (bc[2] * nb + bc[1]) * nb + bc[0];
        if (MODE == 0) {
            atomicAdd(&bb.counts[brick], 1);
        } else {
            const long long pos = (long long)bb.offsets[brick] + atomicAdd(&bb.cursors[brick], 1);
            if (pos < bb.list_cap) bb.list[pos] = q;
            else bb.ctl[3] = 1;
        }
    }
}

// Single-block scan: offsets / item offsets, totals, counter reset.
__global__ void __launch_bounds__(1024) k_bin_scan(BrickGeom g, BinBuffers bb, int K)
{
    const int nbk = g.nbricks;
    const int tid = threadIdx.x;
    const int per = (nbk + 1023) / 1024;
    const int b0 = tid * per, b1 = min(nbk, b0 + per);
    long long sc = 0, si = 0;
    for (int b = b0; b < b1; b++) {
        const int cnt = bb.counts[b];
        sc += cnt;
        si += (cnt + K - 1) / K;
    }
    __shared__ long long s_c[1024], s_i[1024];
    s_c[tid] = sc;
    s_i[tid] = si;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long vc = tid >= off ? s_c[tid - off] : 0;
        long long vi = tid >= off ? s_i[tid - off] : 0;
        __syncthreads();
        s_c[tid] += vc;
        s_i[tid] += vi;
        __syncthreads();
    }
    long long ec = s_c[tid] - sc, ei = s_i[tid] - si;
    for (int b = b0; b < b1; b++) {
        const int cnt = bb.counts[b];
        bb.offsets[b] = (int)ec;
        bb.item_off[b] = (int)ei;
        bb.cursors[b] = 0;
        ec += cnt;
        ei += (cnt + K - 1) / K;
    }
    if (tid == 1023) {
        bb.offsets[nbk] = (int)s_c[1023];
        bb.item_off[nbk] = (int)s_i[1023];
        const long long items = s_i[1023] < bb.item_cap ? s_i[1023] : bb.item_cap;
        bb.ctl[0] = 0;
        bb.ctl[1] = (int)items;
        bb.ctl[2] = (int)(s_c[1023] < bb.list_cap ? s_c[1023] : bb.list_cap);
        if (s_c[1023] > bb.list_cap || s_i[1023] > bb.item_cap) bb.ctl[3] = 1;
    }
}

__global__ void k_bin_items(BrickGeom g, BinBuffers bb, int K)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= g.nbricks) return;
    const int cnt = bb.counts[b];
    const int base = bb.item_off[b];
    const int off = bb.offsets[b];
    for (int i = 0; i * K < cnt; i++) {
        if ((long long)base + i >= bb.item_cap) break;
        WorkItem w;
        w.brick = b;
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
    if ((e = launch_fill_i32(b.counts, 0, g.nbricks, s)) != hipSuccess) return e;
    if ((e = launch_fill_i32(b.ctl, 0, 4, s)) != hipSuccess) return e;
    if (nslices > 0) {
        dim3 grid((g.nb * g.nb + 255) / 256, nslices);
        hipLaunchKernelGGL(k_bin<0>, grid, dim3(256), 0, s, t3d, wq, nslices, c, g, b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, g, b, K);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (nslices > 0) {
        dim3 grid((g.nb * g.nb + 255) / 256, nslices);
        hipLaunchKernelGGL(k_bin<1>, grid, dim3(256), 0, s, t3d, wq, nslices, c, g, b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_bin_items, dim3((g.nbricks + 255) / 256), dim3(256), 0, s, g, b, K);
    return hipGetLastError();
}

// ------------------------------------------------------------ brick insertion
// Fold one unfolded cell's accumulated (value, weight) into the half-volume
// storage exactly as add_complex_at_fast places each contribution.
__device__ __forceinline__ void fold_cell(float2 *data, float *norm, int P, int H, int i, int j, int k, float re,
                                          float im, float w)
{
    float *d = reinterpret_cast<float *>(data);
    const int wy = j < 0 ? P + j : j, wz = k < 0 ? P + k : k;
    const int ny = j <= 0 ? -j : P - j, nz = k <= 0 ? -k : P - k;
    if (j == 0 && k == 0 && (i == 0 || i == H || i == -H)) {
        const int xs = i < 0 ? -i : i;
        gadd(&d[2 * xs], re);
        d[2 * xs + 1] = 0.f;
        gadd(&norm[xs], w);
        return;
    }
    if (i == 0 || i == H) {
        const size_t a = cidx(i, ny, nz, P, H), b = cidx(i, wy, wz, P, H);
        gadd(&d[2 * a], re);
        gadd(&d[2 * a + 1], -im);
        gadd(&d[2 * b], re);
        gadd(&d[2 * b + 1], im);
        gadd(&norm[b], w);
        return;
    }
    if (i == -H) {
        const size_t a = cidx(H, ny, nz, P, H);
        gadd(&d[2 * a], re);
        gadd(&d[2 * a + 1], -im);
        gadd(&d[2 * a], re);
        gadd(&d[2 * a + 1], -im);
        gadd(&norm[a], w);
        return;
    }
    if (i > 0) {
        const size_t b = cidx(i, wy, wz, P, H);
        gadd(&d[2 * b], re);
        gadd(&d[2 * b + 1], im);
        gadd(&norm[b], w);
        return;
    }
    const size_t a = cidx(-i, ny, nz, P, H);
    gadd(&d[2 * a], re);
    gadd(&d[2 * a + 1], -im);
    gadd(&norm[a], w);
}

// LDS accumulator index of brick cell (i,j,k): the row offset is skewed by
// (9j + 11k) (16^3) or (j + 3k) (8^3) so that the cells hit by one pixel
// row -- a discretised 3-D line in any direction -- spread over the banks.
template <int BS>
__device__ __forceinline__ int acc_index(int i, int j, int k)
{
    if (BS == 16) return ((k * 16 + j) << 4) + ((i + 9 * j + 11 * k) & 15);
    return ((k * 8 + j) << 3) + ((i + j + 3 * k) & 7);
}
template <int BS>
__device__ __forceinline__ void acc_cell(int a, int &i, int &j, int &k)
{
    k = a / (BS * BS);
    j = (a / BS) % BS;
    if (BS == 16) i = ((a & 15) - 9 * j - 11 * k) & 15;
    else i = ((a & 7) - j - 3 * k) & 7;
}

template <int BS>
__global__ void __launch_bounds__(512) k_brick_insert(const float2 *__restrict__ slices, int nslices, int nsym,
                                                      const float *__restrict__ t3d, const float *__restrict__ wq,
                                                      InsertConst c, BrickGeom g, BinBuffers bb, float2 *data,
                                                      float *norm, const float *__restrict__ exptab_g,
                                                      const float *__restrict__ rtab_g)
{
    constexpr int NC = BS * BS * BS;
    constexpr int NW = 8;  // waves per block
    __shared__ float acc_re[NC], acc_im[NC], acc_w[NC];
    __shared__ float etab[EXPN];
    __shared__ int s_item;
    extern __shared__ float rtab[];  // P+1 entries: fl(i/P), i in [-H, H]
    const int H = c.H, P = c.P, nb = g.nb;
    const float Pf = c.Pf;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < EXPN; i += blockDim.x) etab[i] = exptab_g[i];
    for (int i = tid; i <= P; i += blockDim.x) rtab[i] = rtab_g[i];
    const int total_items = bb.ctl[1];
    const size_t slice_elems = (size_t)P * (H + 1);

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&bb.ctl[0], 1);
        for (int i = tid; i < NC; i += blockDim.x) {
            acc_re[i] = 0.f;
            acc_im[i] = 0.f;
            acc_w[i] = 0.f;
        }
        __syncthreads();
        const int it = s_item;
        if (it >= total_items) break;
        const WorkItem w = bb.items[it];
        const int bx = w.brick % nb, by = (w.brick / nb) % nb, bz = w.brick / (nb * nb);
        const int lo0 = -H + BS * bx, lo1 = -H + BS * by, lo2 = -H + BS * bz;

        for (int si = wave; si < w.count; si += NW) {
            const int q = __builtin_amdgcn_readfirstlane(bb.list[w.start + si]);
            if (q < 0 || q >= nslices) continue;  // never written list entry
            const float *m = t3d + 9 * (size_t)q;
            float mm[6];
#pragma unroll
            for (int d = 0; d < 6; d++) mm[d] = m[d];
            const float weight = wq[q];
            const float2 *S = slices + (size_t)(q / nsym) * slice_elems;
            // Rows and x intervals of the pixels whose position can fall in
            // [lo-1, lo+BS) (+/- 1e-4 covers the FP32 coordinate error).
            const double L[3] = {lo0 - 1 - 1e-4, lo1 - 1 - 1e-4, lo2 - 1 - 1e-4};
            const double U[3] = {lo0 + BS + 1e-4, lo1 + BS + 1e-4, lo2 + BS + 1e-4};
            double ymin = 0, ymax = 0, xmin = 0, xmax = 0, uu[3], vv[3], iu[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                uu[d] = mm[d];
                vv[d] = mm[3 + d];
                iu[d] = fabs(uu[d]) > 1e-9 ? 1.0 / uu[d] : 0.0;
                ymin += vv[d] >= 0 ? L[d] * vv[d] : U[d] * vv[d];
                ymax += vv[d] >= 0 ? U[d] * vv[d] : L[d] * vv[d];
                xmin += uu[d] >= 0 ? L[d] * uu[d] : U[d] * uu[d];
                xmax += uu[d] >= 0 ? U[d] * uu[d] : L[d] * uu[d];
            }
            const int ya = max(-H, (int)ceil(ymin - 1e-3));
            const int yb = min(H - 1, (int)floor(ymax + 1e-3));
            // (BS+1)*sum|v| + 1 <= 31 rows for BS <= 16; clamp to the wave
            const int nrows = min(64, yb - ya + 1);
            int cnt = 0, xa = 0;
            if (lane < nrows) {
                const int y = ya + lane;
                double xl = fmax(0.0, xmin - 1e-3), xh = fmin((double)H, xmax + 1e-3);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double yv = y * vv[d];
                    if (iu[d] != 0.0) {
                        const double t1 = (L[d] - yv) * iu[d], t2 = (U[d] - yv) * iu[d];
                        xl = fmax(xl, fmin(t1, t2) - 1e-3);
                        xh = fmin(xh, fmax(t1, t2) + 1e-3);
                    } else if (yv < L[d] || yv > U[d]) {
                        xh = -1.0;
                    }
                }
                if (!c.corners) xh = fmin(xh, floor(sqrt((double)(H * H + H - y * y))));
                xa = (int)ceil(xl);
                const int xb = (int)floor(xh);
                cnt = xb >= xa ? xb - xa + 1 : 0;
            }
            // inclusive wave scan of cnt
            int incl = cnt;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            const int total = __shfl(incl, 63, 64);
            const int excl = incl - cnt;
            for (int base = 0; base < total; base += 64) {
                const int gidx = base + lane;
                // rows overlapping the window [base, base+64) (a contiguous run)
                const unsigned long long ov =
                    __ballot(lane < nrows && cnt > 0 && excl < base + 64 && excl + cnt > base);
                const int r_lo = __ffsll((long long)ov) - 1;
                const int r_hi = 63 - __clzll((long long)ov);
                // row of slot gidx: the last overlapping row r with excl_r <= gidx
                int rrow = 0, rex = 0, rxa = 0;
                for (int r = r_lo; r <= r_hi; r++) {
                    const int er = __builtin_amdgcn_readlane(excl, r);
                    const int xr = __builtin_amdgcn_readlane(xa, r);
                    const bool sel = er <= gidx;
                    rrow = sel ? r : rrow;
                    rex = sel ? er : rex;
                    rxa = sel ? xr : rxa;
                }
                if (gidx >= total) continue;
                const int y = ya + rrow;
                const int x = rxa + (gidx - rex);
                const float2 dt = slice_value(S, x, y, P, H);  // issued before the coordinate math
                bool ok = c.corners || (x * x + y * y <= H * H + H);
                float xx, yy, zz;
                pixel_coords(x, y, mm, rtab, H, Pf, xx, yy, zz);
                const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
                ok = ok && !(x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H);
                ok = ok && x0 >= lo0 - 1 && x0 <= lo0 + BS - 1 && y0 >= lo1 - 1 && y0 <= lo1 + BS - 1 &&
                     z0 >= lo2 - 1 && z0 <= lo2 + BS - 1;
                if (!ok) continue;
#pragma unroll
                for (int dk = 0; dk < 2; dk++) {
                    const int k = z0 + dk;
                    if (k < lo2 || k >= lo2 + BS || k > H) continue;
#pragma unroll
                    for (int dj = 0; dj < 2; dj++) {
                        const int j = y0 + dj;
                        if (j < lo1 || j >= lo1 + BS || j > H) continue;
#pragma unroll
                        for (int di = 0; di < 2; di++) {
                            const int i = x0 + di;
                            if (i < lo0 || i >= lo0 + BS || i > H) continue;
                            const float gg = corner_weight(c.mode, i, j, k, xx, yy, zz, c.h, weight, etab);
                            const int a = acc_index<BS>(i - lo0, j - lo1, k - lo2);
                            atomicAdd(&acc_re[a], dt.x * gg);
                            atomicAdd(&acc_im[a], dt.y * gg);
                            atomicAdd(&acc_w[a], gg);
                        }
                    }
                }
            }
        }
        __syncthreads();
        for (int a = tid; a < NC; a += blockDim.x) {
            const float re = acc_re[a], im = acc_im[a], wv = acc_w[a];
            if (re == 0.f && im == 0.f && wv == 0.f) continue;
            int ci, cj, ck;
            acc_cell<BS>(a, ci, cj, ck);
            const int i = lo0 + ci, j = lo1 + cj, k = lo2 + ck;
            if (i > H || j > H || k > H) continue;
            fold_cell(data, norm, P, H, i, j, k, re, im, wv);
        }
        __syncthreads();
    }
}

hipError_t launch_brick_insert(const float2 *slices, int nslices, int nsym, const float *t3d,
                               const float *wq, InsertConst c, BrickGeom g, BinBuffers b, float2 *data, float *norm,
                               const float *exptab, const float *rtab, int grid, hipStream_t s)
{
    const size_t lds = sizeof(float) * (c.P + 1);
    if (g.Bs == 16)
        hipLaunchKernelGGL(k_brick_insert<16>, dim3(grid), dim3(512), lds, s, slices, nslices, nsym, t3d, wq, c, g,
                           b, data, norm, exptab, rtab);
    else
        hipLaunchKernelGGL(k_brick_insert<8>, dim3(grid), dim3(512), lds, s, slices, nslices, nsym, t3d, wq, c, g,
                           b, data, norm, exptab, rtab);
    return hipGetLastError();
}

}  // namespace fh

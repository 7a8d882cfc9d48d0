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
//  k_column_insert   one wave per work item: each lane owns one of the
//                    brick's 8x8 voxel columns along the item's axis and
//                    gathers the pixels whose stencil reaches that column;
//                    accumulation is lane-private (no atomics).  Each item
//                    writes its partial brick to HBM.
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
// column along the dominant normal axis at most 5 bricks qualify.  A slice
// is binned under key 3 * brick + (its dominant normal axis), so that every
// work item's slices share one column direction.
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
    const int nb = g.nb, BS = BRICK, H = c.H;
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
        const int key = 3 * ((bc[2] * nb + bc[1]) * nb + bc[0]) + f.a;
        if (MODE == 0) {
            atomicAdd(&bb.counts[key], 1);
        } else {
            const long long pos = (long long)bb.offsets[key] + atomicAdd(&bb.cursors[key], 1);
            if (pos < bb.list_cap) bb.list[pos] = q;
            else bb.ctl[3] = 1;
        }
    }
}

// Single-block scan: offsets / item offsets, totals, counter reset.
__global__ void __launch_bounds__(1024) k_bin_scan(BrickGeom g, BinBuffers bb, int K)
{
    const int nbk = g.nkeys;
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
    if (b >= g.nkeys) return;
    const int cnt = bb.counts[b];
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
    if ((e = launch_fill_i32(b.counts, 0, g.nkeys, s)) != hipSuccess) return e;
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
    hipLaunchKernelGGL(k_bin_items, dim3((g.nkeys + 255) / 256), dim3(256), 0, s, g, b, K);
    return hipGetLastError();
}

// ------------------------------------------------------------ column-gather insertion
// A work item is (brick, axis A, up to K slices whose dominant normal axis is
// A).  Lane l of the item's wave owns the brick's voxel column
// (c1, c2) = (lo[B1] + l%8, lo[B2] + l/8) along A (B1 < B2 the other two
// axes).  A pixel at p adds to cells floor(p) + {0,1}^3, so it reaches the
// column iff p[B] lies in [c_B - 1, c_B + 1) on both B axes: per slice the
// lane visits that parallelogram of slice pixels and adds each pixel's two
// cells (floor(p[A]) and +1) of its column.  Every (pixel, cell)
// contribution of the reference loop lands in exactly one (brick, column)
// -- the accumulation needs no atomics.
//
// Per-wave LDS accumulator: [field re/im/w][slot][lane], slot s holding the
// column cell lo[A] - 1 + s (s = 0 and CSLOT-1 are pads for stencil cells
// just outside the brick, never merged).  Lanes of one access hit distinct
// lane columns, so the read-modify-writes are bank-conflict free.
constexpr int CSLOT = BRICK + 2;
constexpr int CACC = 3 * CSLOT * 64;  // floats of one wave's accumulator
#define COL_DLT 2e-3f                 // margin on p[B]: FP32 position error is < 3e-4 for P <= 1024
#define COL_EPS 1e-3f                 // margin on enumeration bounds

// Sweep of one slice over the wave's 64 columns.  The parallelogram is
// enumerated in rows of the slice coordinate (x or y) with fewer rows,
// relative to an integer pixel (xc0, yc0) near the brick's preimage so that
// the bounds are small numbers; the bounds carry margins and every pixel is
// then checked with the reference's own FP32 coordinates.
template <int A>
__device__ __forceinline__ void column_slice(const float2 *__restrict__ S, const float mm[6], float weight,
                                             const InsertConst &c, const int lo[3], int lane, float *acc,
                                             const float *etab, const float *rtab)
{
    constexpr int B1 = A == 0 ? 1 : 0, B2 = A == 2 ? 1 : 2;
    const int H = c.H, P = c.P;
    const int c1 = lo[B1] + (lane & 7), c2 = lo[B2] + (lane >> 3);
    // slice set-up (wave-uniform)
    const float u1 = mm[B1], u2 = mm[B2], v1 = mm[3 + B1], v2 = mm[3 + B2];
    const float D = u1 * v2 - u2 * v1;  // = +-n[A], |n[A]| >= 1/sqrt(3)
    if (!(fabsf(D) >= 1e-3f)) return;   // not a rotation (or NaN): keeps every bound finite
    const float iD = 1.0f / D;
    const float C1 = (float)lo[B1] + 3.5f, C2 = (float)lo[B2] + 3.5f;
    const int xc0 = (int)rintf((v2 * C1 - v1 * C2) * iD), yc0 = (int)rintf((u1 * C2 - u2 * C1) * iD);
    const float P01 = (float)xc0 * u1 + (float)yc0 * v1, P02 = (float)xc0 * u2 + (float)yc0 * v2;
    const bool rows_y = fabsf(u1) + fabsf(u2) <= fabsf(v1) + fabsf(v2);
    // p[B] = S * a_B + R * b_B (+ P0_B): R the row coordinate, S the in-row one
    const float a1 = rows_y ? u1 : v1, a2 = rows_y ? u2 : v2, b1 = rows_y ? v1 : u1, b2 = rows_y ? v2 : u2;
    const float iDp = rows_y ? iD : -iD;  // 1 / (a1 b2 - a2 b1)
    const float rr = (fabsf(a1) + fabsf(a2)) * (1.0f + COL_DLT) * fabsf(iDp);
    // in-row bounds from each B axis whose coefficient is not small; one of
    // |a1|, |a2| is >= 0.408 because |det| >= 0.577 and |(b1, b2)| <= 1
    const bool use1 = fabsf(a1) >= 0.25f, use2 = fabsf(a2) >= 0.25f;
    const float ia1 = use1 ? 1.0f / a1 : 0.f, ia2 = use2 ? 1.0f / a2 : 0.f;
    const float s1 = b1 * ia1, s2 = b2 * ia2;
    // lane set-up
    const float e1 = (float)c1 - P01, e2 = (float)c2 - P02;
    const float rmid = (a1 * e2 - a2 * e1) * iDp;
    int Ra = (int)ceilf(rmid - rr - COL_EPS), Rb = (int)floorf(rmid + rr + COL_EPS);
    int Smin, Smax;
    if (rows_y) {
        Ra = max(Ra, -H - yc0);
        Rb = min(Rb, H - 1 - yc0);
        Smin = -xc0;
        Smax = H - xc0;
    } else {
        Ra = max(Ra, -xc0);
        Rb = min(Rb, H - xc0);
        Smin = -H - yc0;
        Smax = H - 1 - yc0;
    }
    if (c1 > H || c2 > H) Rb = Ra - 1;  // column outside the unfolded grid
    const float k10 = (e1 - (1.0f + COL_DLT)) * ia1, k11 = (e1 + (1.0f + COL_DLT)) * ia1;
    const float k20 = (e2 - (1.0f + COL_DLT)) * ia2, k21 = (e2 + (1.0f + COL_DLT)) * ia2;
    int R = Ra - 1, Sc = 1, Sb = 0;
    for (;;) {
        while (Sc > Sb) {  // next non-empty row
            if (++R > Rb) break;
            const float Rf = (float)R;
            float lo_s = (float)Smin, hi_s = (float)Smax;
            if (use1) {
                const float t0 = k10 - Rf * s1, t1 = k11 - Rf * s1;
                lo_s = fmaxf(lo_s, fminf(t0, t1));
                hi_s = fminf(hi_s, fmaxf(t0, t1));
            }
            if (use2) {
                const float t0 = k20 - Rf * s2, t1 = k21 - Rf * s2;
                lo_s = fmaxf(lo_s, fminf(t0, t1));
                hi_s = fminf(hi_s, fmaxf(t0, t1));
            }
            Sc = max(Smin, (int)ceilf(lo_s - COL_EPS));
            Sb = min(Smax, (int)floorf(hi_s + COL_EPS));
        }
        if (R > Rb) break;
        const int x = rows_y ? xc0 + Sc : xc0 + R;
        const int y = rows_y ? yc0 + R : yc0 + Sc;
        Sc++;
        if (!c.corners && x * x + y * y > H * H + H) continue;  // hypot_fast_int(x,y) > P/2
        const float2 dt = slice_value(S, x, y, P, H);
        float xx, yy, zz;
        pixel_coords(x, y, mm, rtab, H, c.Pf, xx, yy, zz);
        const int f0 = (int)floorf(xx), f1 = (int)floorf(yy), f2 = (int)floorf(zz);
        if (f0 < -H - 1 || f1 < -H - 1 || f2 < -H - 1 || f0 > H || f1 > H || f2 > H) continue;
        const int fa = A == 0 ? f0 : (A == 1 ? f1 : f2);
        const int fb1 = B1 == 0 ? f0 : f1;
        const int fb2 = B2 == 1 ? f1 : f2;
        if ((unsigned)(c1 - fb1) > 1u || (unsigned)(c2 - fb2) > 1u) continue;
        const int t = fa - lo[A];
        if (t < -1 || t > BRICK - 1) continue;
#pragma unroll
        for (int dk = 0; dk < 2; dk++) {
            const int ca = fa + dk;
            const int i = A == 0 ? ca : c1;
            const int j = A == 0 ? c1 : (A == 1 ? ca : c2);
            const int k = A == 2 ? ca : c2;
            const float gg = corner_weight(c.mode, i, j, k, xx, yy, zz, c.h, weight, etab);
            float *a = acc + (t + 1 + dk) * 64 + lane;
            a[0] += dt.x * gg;
            a[CSLOT * 64] += dt.y * gg;
            a[2 * CSLOT * 64] += gg;
        }
    }
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Partial brick layout: [field][z][y][x] of the brick's local cells.
template <int A>
__device__ __forceinline__ void column_flush(const float *acc, float *out, int lane)
{
#pragma unroll
    for (int w = 0; w < BRICK; w++) {
        const int cc = lane + 64 * w;
        const int lx = cc & 7, ly = (cc >> 3) & 7, lz = cc >> 6;
        const int slot = (A == 0 ? lx : (A == 1 ? ly : lz)) + 1;
        const int ln = A == 0 ? ly + 8 * lz : (A == 1 ? lx + 8 * lz : lx + 8 * ly);
        const float *a = acc + slot * 64 + ln;
        out[cc] = a[0];
        out[512 + cc] = a[CSLOT * 64];
        out[1024 + cc] = a[2 * CSLOT * 64];
    }
}

template <int A>
__device__ __forceinline__ void column_item(const float2 *__restrict__ slices, int nslices, int nsym,
                                            const float *__restrict__ t3d, const float *__restrict__ wq,
                                            const InsertConst &c, const int *__restrict__ list, const WorkItem &w,
                                            const int lo[3], int lane, float *acc, const float *etab,
                                            const float *rtab, float *out)
{
    const size_t slice_elems = (size_t)c.P * (c.H + 1);
    for (int si = 0; si < w.count; si++) {
        const int q = __builtin_amdgcn_readfirstlane(list[w.start + si]);
        if (q < 0 || q >= nslices) continue;  // never written list entry
        const float *m = t3d + 9 * (size_t)q;
        float mm[6];
#pragma unroll
        for (int d = 0; d < 6; d++) mm[d] = m[d];
        column_slice<A>(slices + (size_t)(q / nsym) * slice_elems, mm, wq[q], c, lo, lane, acc, etab, rtab);
    }
    wave_sync_lds();
    column_flush<A>(acc, out, lane);
    wave_sync_lds();
}

__global__ void __launch_bounds__(256) k_column_insert(const float2 *__restrict__ slices, int nslices, int nsym,
                                                       const float *__restrict__ t3d, const float *__restrict__ wq,
                                                       InsertConst c, BrickGeom g, BinBuffers bb,
                                                       const float *__restrict__ exptab_g,
                                                       const float *__restrict__ rtab_g)
{
    extern __shared__ float lds[];
    const int rlen = (c.P + 2 + 3) & ~3;
    float *etab = lds;
    float *rtab = lds + EXPN;  // P+2 entries: fl(i/P), i in [-H, H+1]
    const int lane = threadIdx.x & 63;
    float *acc = lds + EXPN + rlen + (threadIdx.x >> 6) * CACC;
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
        const int lo[3] = {-H + BRICK * (brick % nb), -H + BRICK * ((brick / nb) % nb), -H + BRICK * (brick / (nb * nb))};
        for (int s = 0; s < 3 * CSLOT; s++) acc[s * 64 + lane] = 0.f;
        float *out = bb.part + (size_t)it * PART_FLOATS;
        if (A == 0) column_item<0>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, etab, rtab, out);
        else if (A == 1) column_item<1>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, etab, rtab, out);
        else column_item<2>(slices, nslices, nsym, t3d, wq, c, bb.list, w, lo, lane, acc, etab, rtab, out);
    }
}

hipError_t launch_column_insert(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                                InsertConst c, BrickGeom g, BinBuffers b, const float *exptab, const float *rtab,
                                int grid, int waves_per_block, hipStream_t s)
{
    if (waves_per_block < 1 || waves_per_block > 4) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (EXPN + ((c.P + 2 + 3) & ~3) + (size_t)waves_per_block * CACC);
    hipLaunchKernelGGL(k_column_insert, dim3(grid), dim3(64 * waves_per_block), lds, s, slices, nslices, nsym, t3d,
                       wq, c, g, b, exptab, rtab);
    return hipGetLastError();
}

// ------------------------------------------------------------ merge of partial bricks
// Sum over the work items of the brick holding unfolded cell (i, j, k).
__device__ __forceinline__ bool partial_sum(const float *__restrict__ part, const int *__restrict__ item_off, int nb,
                                            int H, int i, int j, int k, float &re, float &im, float &w)
{
    const int ui = i + H, uj = j + H, uk = k + H;
    const int b = ((uk >> 3) * nb + (uj >> 3)) * nb + (ui >> 3);
    const int cc = ((uk & 7) << 6) | ((uj & 7) << 3) | (ui & 7);
    const int i0 = item_off[3 * b], i1 = item_off[3 * b + 3];
    re = im = w = 0.f;
    for (int it = i0; it < i1; it++) {
        const float *p = part + (size_t)it * PART_FLOATS + cc;
        re += p[0];
        im += p[512];
        w += p[1024];
    }
    return i1 > i0;
}

// Storage voxel (xs, ys, zs) of the half volume receives, per
// add_complex_at_fast (libEM/emdata_core.h:368-413): the direct cells
// (xs, j, k) with wrap(j) = ys, wrap(k) = zs; the conjugate cells (-xs, j, k)
// with wrapneg(j) = ys, wrapneg(k) = zs; on the x = 0 and x = P/2 planes the
// conjugate copies of the double insertion (value only) and, at x = P/2,
// the x = -P/2 cells added twice with their weight once; at (0,0,0) and
// (P/2,0,0) the real part and weight with the imaginary part set to 0.
__global__ void __launch_bounds__(256) k_merge_partials(const float *__restrict__ part, const int *__restrict__ item_off,
                                                        BrickGeom g, InsertConst c, float2 *data, float *norm)
{
    const int H = c.H, P = c.P, nb = g.nb;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)P * P * (H + 1)) return;
    const int xs = (int)(idx % (H + 1));
    const long long rest = idx / (H + 1);
    const int ys = (int)(rest % P), zs = (int)(rest / P);
    float r, i, w;
    if (ys == 0 && zs == 0 && (xs == 0 || xs == H)) {
        partial_sum(part, item_off, nb, H, xs, 0, 0, r, i, w);
        if (xs == H) {
            float r2, i2, w2;
            partial_sum(part, item_off, nb, H, -H, 0, 0, r2, i2, w2);
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
            partial_sum(part, item_off, nb, H, xs, jd[a], kd[b], r, i, w);
            dre += r;
            dim += i;
            dw += w;
        }
    for (int a = 0; a < njn; a++)
        for (int b = 0; b < nkn; b++) {
            if (xs == 0 || xs == H) {
                partial_sum(part, item_off, nb, H, xs, jn[a], kn[b], r, i, w);
                dre += r;
                dim -= i;
                if (xs == H) {
                    partial_sum(part, item_off, nb, H, -H, jn[a], kn[b], r, i, w);
                    dre += 2.f * r;
                    dim -= 2.f * i;
                    dw += w;
                }
            } else {
                partial_sum(part, item_off, nb, H, -xs, jn[a], kn[b], r, i, w);
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
    hipLaunchKernelGGL(k_merge_partials, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b.part, b.item_off, g, c,
                       data, norm);
    return hipGetLastError();
}

}  // namespace fh

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

template <int BS>
__global__ void __launch_bounds__(256) k_brick_insert(const float2 *__restrict__ slices, int nslices, int nsym,
                                                      const float *__restrict__ t3d, const float *__restrict__ wq,
                                                      InsertConst c, BrickGeom g, BinBuffers bb, float2 *data,
                                                      float *norm, const float *__restrict__ exptab_g,
                                                      const float *__restrict__ rtab_g)
{
    constexpr int NC = BS * BS * BS;
    constexpr int NW = 4;  // waves per block
    __shared__ float acc_re[NC], acc_im[NC], acc_w[NC];
    __shared__ float etab[EXPN];
    __shared__ int s_excl[NW][64], s_xa[NW][64];
    __shared__ int s_item;
    extern __shared__ float rtab[];  // P+1 entries
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
        const int lo[3] = {-H + BS * bx, -H + BS * by, -H + BS * bz};

        for (int si = wave; si < w.count; si += NW) {
            const int q = __builtin_amdgcn_readfirstlane(bb.list[w.start + si]);
            if (q < 0 || q >= nslices) continue;  // never written list entry
            const float *m = t3d + 9 * (size_t)q;
            float mm[6];
#pragma unroll
            for (int d = 0; d < 6; d++) mm[d] = m[d];
            const float weight = wq[q];
            const float2 *S = slices + (size_t)(q / nsym) * slice_elems;
            // rows and x intervals of pixels whose position can fall in
            // [lo-1, lo+BS) (+/- 1e-4 covers the FP32 coordinate error)
            double L[3], U[3], uu[3], vv[3];
            double ymin = 0, ymax = 0, xmin = 0, xmax = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                L[d] = lo[d] - 1 - 1e-4;
                U[d] = lo[d] + BS + 1e-4;
                uu[d] = mm[d];
                vv[d] = mm[3 + d];
                ymin += vv[d] >= 0 ? L[d] * vv[d] : U[d] * vv[d];
                ymax += vv[d] >= 0 ? U[d] * vv[d] : L[d] * vv[d];
                xmin += uu[d] >= 0 ? L[d] * uu[d] : U[d] * uu[d];
                xmax += uu[d] >= 0 ? U[d] * uu[d] : L[d] * uu[d];
            }
            const int ya = max(-H, (int)ceil(ymin - 1e-3));
            const int yb = min(H - 1, (int)floor(ymax + 1e-3));
            // (BS+1)*sum|v| + 1 <= 31 rows for BS <= 16; clamp to the 64-lane table
            const int nrows = min(64, yb - ya + 1);
            int cnt = 0, xa = 0;
            if (lane < nrows) {
                const int y = ya + lane;
                double xl = fmax(0.0, xmin - 1e-3), xh = fmin((double)H, xmax + 1e-3);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const double yv = y * vv[d];
                    if (fabs(uu[d]) > 1e-9) {
                        const double t1 = (L[d] - yv) / uu[d], t2 = (U[d] - yv) / uu[d];
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
            s_excl[wave][lane] = incl - cnt;
            s_xa[wave][lane] = xa;
            __builtin_amdgcn_wave_barrier();
            for (int base = 0; base < total; base += 64) {
                const int gidx = base + lane;
                if (gidx < total) {
                    // largest row r with excl[r] <= gidx
                    int lo_r = 0, hi_r = nrows - 1;
                    while (lo_r < hi_r) {
                        const int mid = (lo_r + hi_r + 1) >> 1;
                        if (s_excl[wave][mid] <= gidx) lo_r = mid;
                        else hi_r = mid - 1;
                    }
                    const int y = ya + lo_r;
                    const int x = s_xa[wave][lo_r] + (gidx - s_excl[wave][lo_r]);
                    bool ok = c.corners || (x * x + y * y <= H * H + H);
                    float xx, yy, zz;
                    pixel_coords(x, y, mm, rtab, H, Pf, xx, yy, zz);
                    const int x0 = (int)floorf(xx), y0 = (int)floorf(yy), z0 = (int)floorf(zz);
                    ok = ok && !(x0 < -H - 1 || y0 < -H - 1 || z0 < -H - 1 || x0 > H || y0 > H || z0 > H);
                    ok = ok && x0 >= lo[0] - 1 && x0 <= lo[0] + BS - 1 && y0 >= lo[1] - 1 &&
                         y0 <= lo[1] + BS - 1 && z0 >= lo[2] - 1 && z0 <= lo[2] + BS - 1;
                    if (ok) {
                        const float2 dt = slice_value(S, x, y, P, H);
#pragma unroll
                        for (int dk = 0; dk < 2; dk++) {
                            const int k = z0 + dk;
                            if (k < lo[2] || k >= lo[2] + BS || k > H) continue;
#pragma unroll
                            for (int dj = 0; dj < 2; dj++) {
                                const int j = y0 + dj;
                                if (j < lo[1] || j >= lo[1] + BS || j > H) continue;
#pragma unroll
                                for (int di = 0; di < 2; di++) {
                                    const int i = x0 + di;
                                    if (i < lo[0] || i >= lo[0] + BS || i > H) continue;
                                    const float gg = corner_weight(c.mode, i, j, k, xx, yy, zz, c.h, weight, etab);
                                    const int cell = ((k - lo[2]) * BS + (j - lo[1])) * BS + (i - lo[0]);
                                    atomicAdd(&acc_re[cell], dt.x * gg);
                                    atomicAdd(&acc_im[cell], dt.y * gg);
                                    atomicAdd(&acc_w[cell], gg);
                                }
                            }
                        }
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        for (int cell = tid; cell < NC; cell += blockDim.x) {
            const float re = acc_re[cell], im = acc_im[cell], wv = acc_w[cell];
            if (re == 0.f && im == 0.f && wv == 0.f) continue;
            const int i = lo[0] + cell % BS, j = lo[1] + (cell / BS) % BS, k = lo[2] + cell / (BS * BS);
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
        hipLaunchKernelGGL(k_brick_insert<16>, dim3(grid), dim3(256), lds, s, slices, nslices, nsym, t3d, wq, c, g,
                           b, data, norm, exptab, rtab);
    else
        hipLaunchKernelGGL(k_brick_insert<8>, dim3(grid), dim3(256), lds, s, slices, nslices, nsym, t3d, wq, c, g,
                           b, data, norm, exptab, rtab);
    return hipGetLastError();
}

}  // namespace fh

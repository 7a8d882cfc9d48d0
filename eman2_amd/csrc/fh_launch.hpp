// Host-side declarations of the fourier_hip kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

// Mixed-radix plan for the LDS Stockham FFT: N = prod(radix[i]), radices
// in {2,3,4,5,6,8}; pass p has stride ns[p] = prod(radix[<p]), mag[p] =
// ceil(2^32 / ns[p]) (exact j / ns for j < 2^16) and twiddle step
// twm[p] = N / (ns[p] * radix[p]).
struct FftPlan {
    int N;
    int npass;
    int radix[16];
    int ns[16];
    unsigned mag[16];
    int twm[16];
};
bool make_plan(int N, FftPlan *p);

// CTF_SIGN constants (libEM/ctf.cpp:985-992) computed on the host.
struct CtfConst {
    float g1, g2, acac, ds;
};

// Per-volume constants of the inserters.
struct InsertConst {
    int P;       // padded box
    int H;       // P/2
    float Pf;    // (float)P
    float h;     // 1.0f / EMConsts::I2G
    int mode;    // 0 gauss_2, 1 trilinear, 2 gauss_var, 3 gauss_5, 4 nearest_neighbor, 5 gauss_3, 6 kaiser_bessel,
                 // 7 gridding_5, 8 gridding_7, 9 kaiser_bessel_derived; 10 fourier_iter, 11 fourier_iter with
                 // a reference volume
    int corners;
    int rcut;            // slice pixels to round(hypot(x, y)) <= rcut (P/2; P/2 + 3 for fourier_iter)
    const float2 *ref;   // fourier_iter's reference volume (mode 11), else null
};

// Row-major 3x3 rotation passed by value.
struct Mat9 {
    float m[9];
};

// Work item of the batched insertion: slices list[start, start+count) of
// bin key `key` = 3 * brick + dominant normal axis of the slices.
struct WorkItem {
    int key;
    int start;
    int count;
    int pad;
};

// Per-slice plane record of the binning, grouped by dominant normal axis a:
// t3d rows u, v and the column slopes r1 = -n[b1]/n[a], r2 = -n[b2]/n[a]
// (n = u x v in double), so the plane's a coordinate is r1 p_b1 + r2 p_b2.
struct SliceRec {
    float u[3], v[3];
    double r1, r2;
    int q;    // slice index
    int pad;
};

// Device buffers of the binning stage and the partial bricks.
struct BinBuffers {
    int *counts;      // nkeys
    int *offsets;     // nkeys + 1
    int *item_off;    // nkeys + 1: first work item of each key
    int *tile;        // 2 per scan tile of 1024 keys: count and item sums / offsets
    int *list;        // capacity
    WorkItem *items;  // item capacity
    int *ctl;         // [0] work counter, [1] total items, [2] total pairs, [3] overflow flag,
                      // [4..6] slices per dominant axis
    SliceRec *recs;   // 3 segments of rec_cap: the slices of each dominant axis
    long long rec_cap;
    float *part;      // partial bricks, BrickGeom::part_floats per work item
    long long list_cap;
    long long item_cap;
    long long part_cap;  // work items the partial buffer holds
};

constexpr int BRICK = 8;                           // brick edge, in stencil origins
constexpr int HALO = BRICK + 1;                    // cells a brick's origins reach per axis
constexpr int HALO_CELLS = HALO * HALO * HALO;     // cells of one partial brick
constexpr int PART_FLOATS = 3 * HALO_CELLS;        // re, im, weight of one partial brick

// Bricks over the stencil origins: brick b owns origins [ow + bs*b, ow + bs*(b+1))
// per axis, i.e. the pixel positions p with origin(p) in that range, which is
// p in [po + bs*b, po + bs*(b+1)) (po = ow + the origin's offset: origin(p) =
// floor(p - off)); its partial covers the halo^3 cells [ow + bs*b, ... + halo).
//   gauss_2 / trilinear: origin floor(p), bs 8, halo 9, ow = po = -P/2-1.
//   wide stencils of S cells per axis: bs 4, halo 3 + S (wide_geom).
struct BrickGeom {
    int nb;           // bricks per axis
    int nbricks;      // nb^3
    int nkeys;        // 3 * nbricks
    int bs;           // brick edge, in stencil origins
    int halo;         // cells per axis of a partial brick
    int ow;           // lowest stencil origin of brick 0
    int part_floats;  // 3 * halo^3: re, im, weight of one partial brick
    float po;         // p-space start of brick 0
    int per_brick;    // 1: work items span a brick's three axis keys (the wide-stencil scatter)
};

constexpr int WBS = 4;  // brick edge of the wide-stencil scatter

// stencil cells per axis of the inserters the wide-stencil scatter takes
// (0: not taken); modes as InsertConst::mode
__host__ __device__ constexpr int wide_cells(int mode)
{
    return (mode == 2 || mode == 3 || mode == 10 || mode == 11) ? 6 : mode == 5 ? 3 : (mode == 6 || mode == 7) ? 5
         : mode == 8 ? 7 : mode == 9 ? 9 : 0;
}

// ---- prep (prep_kernels.hip) ----
hipError_t launch_edge_mean(const float *imgs, int n, int B, float *means, hipStream_t s);
hipError_t launch_prep_rows(const float *imgs, int n, int B, int P, const float *means, float2 *slices,
                            const FftPlan &plan, const float2 *tw, hipStream_t s);
hipError_t launch_prep_cols(int n, int B, int P, float2 *slices, const float *defocus, const float *defocus2,
                            CtfConst cc, float scale, const FftPlan &plan, const float2 *tw, hipStream_t s);
hipError_t launch_ift2d_batch(float2 *data, int n, int P, float *out, float scale, const FftPlan &plan,
                              const float2 *tw, hipStream_t s);
int prep_inter_pitch(int B);
bool prep_column_first(int P);
hipError_t launch_prep2d(const float *imgs, int n, int B, int P, const float *means, float2 *slices,
                         float2 *inter, int inter_n, const float *defocus, const float *defocus2, CtfConst cc,
                         float scale, const FftPlan &plan, const float2 *tw, hipStream_t s);
// ---- finish (prep_kernels.hip) ----
hipError_t launch_normalize(float2 *data, float *norm, int P, int sqrtnorm, float *plane_tmp,
                            hipStream_t s);
hipError_t launch_fft3d_inverse_zy(float2 *data, int P, const FftPlan &plan, const float2 *tw,
                                   hipStream_t s);
hipError_t launch_c2r_x(const float2 *data, int P, float *out, float scale, const FftPlan &plan,
                        const float2 *tw, hipStream_t s);
hipError_t launch_fft3d_forward_corner(const float *vol, int P, const float *zeros, float2 *out, const FftPlan &plan,
                                       const float2 *tw, hipStream_t s);
// gauss_fft projector (modes 0-2) of n projections (rec: 12 floats each) from the map's transform pv
hipError_t launch_gauss_fft(const float2 *pv, int P, int n, int mode, float gw, const float *rec, int ramp,
                            float2 *out, hipStream_t s);
hipError_t launch_gausskernelfix2d(float *img, int P, int n, float gauss_width, hipStream_t s);
hipError_t launch_merge_add(float2 *data, float *norm, const float2 *threed, const float *normv, size_t n,
                            hipStream_t s);
hipError_t launch_merge_scale(float2 *data, const float *norm, size_t n, hipStream_t s);
hipError_t launch_fill_f32(float *p, float v, size_t n, hipStream_t s);
hipError_t launch_fill_i32(int *p, int v, size_t n, hipStream_t s);
hipError_t launch_seed_weights(float *norm, int P, float w, hipStream_t s);
hipError_t launch_scale_f32(float *p, float v, size_t n, hipStream_t s);

// ---- insertion (insert_kernels.hip) ----
// preprocess_slice's 2-D resampling of n padded particles (xf: n x 6 inverse rows) -> n x P x P
hipError_t launch_xform2d(const float *imgs, int n, int B, int P, const float *means, const float *xf, float *out,
                          hipStream_t s);
// projection(): the (P, P/2+1) complex slice of rotation rot9 (3x3 rows), shifted by
// (xshift, yshift); center: phase origin to centre for the real-space output
hipError_t launch_project(const float2 *data, const float *norm, InsertConst c, const float *rot9, const float *rtab,
                          float xshift, float yshift, int center, float2 *out, hipStream_t s);
// 2-D do_ift of one complex image in place (data is overwritten) into P x P floats
hipError_t launch_ift2d(float2 *data, int P, float *out, const FftPlan &plan, const float2 *tw, hipStream_t s);
// every slice of a batch through the plain global-atomic insertion (any mode)
hipError_t launch_insert_plain_batch(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                                     InsertConst c, float2 *data, float *norm, const float *exptab, const float *rtab,
                                     const float *rw, hipStream_t s);
hipError_t launch_compose(const float *rots, const float *weights, int n, const float *syms, int nsym,
                          float *t3d, float *wq, hipStream_t s);
// rw: optional radial weights ((P/2+1) floats, class_ssnr weighting) replacing weight
hipError_t launch_insert_plain(const float2 *slice, InsertConst c, const float *t3d9, float weight,
                               float2 *data, float *norm, const float *exptab, const float *rtab,
                               const float *rw, hipStream_t s);
hipError_t launch_bin(const float *t3d, const float *wq, int nslices, InsertConst c, BrickGeom g,
                      BinBuffers b, int slices_per_item, hipStream_t s);
int scan_tiles(const BrickGeom &g);
size_t scatter_lds_bytes(int P, int waves_per_block);
hipError_t launch_scatter_insert(const float2 *slices, int nslices, int nsym, const float *t3d,
                                 const float *wq, InsertConst c, BrickGeom g, BinBuffers b,
                                 const float *exptab, const float *rtab, const float *rw, int grid,
                                 int waves_per_block, hipStream_t s);
hipError_t launch_merge_partials(BrickGeom g, BinBuffers b, InsertConst c, float2 *data, float *norm,
                                 hipStream_t s);
// the wide-stencil modes (wide_cells(mode) > 0) on the bricks of wide_geom
size_t scatter_wide_lds_bytes(int mode, int waves_per_block);
hipError_t launch_scatter_wide(const float2 *slices, int nslices, int nsym, const float *t3d, const float *wq,
                               InsertConst c, BrickGeom g, BinBuffers b, const float *exptab, const float *rtab,
                               const float *rw, int grid, int waves_per_block, hipStream_t s);
int agree_blocks(int P);
hipError_t launch_agree_batch(const float2 *slices, int nsym, const float *t3d, const float *wq, int nslices,
                              InsertConst c, float rmax, int sub, const float2 *data, const float *norm,
                              const float *exptab, const float *rtab, double *acc, hipStream_t s);
hipError_t launch_agree(const float2 *slices, int nsym, const float *t3d, int nslices, InsertConst c, float rmax,
                        const float2 *data, const float *norm, const float *rtab, double *partial, hipStream_t s);

}  // namespace fh

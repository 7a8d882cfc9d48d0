/*
 * fourier_oracle.h -- CPU restatement of EMAN2's direct Fourier inversion
 * ("fourier" Reconstructor, gauss_2 / trilinear inserters) used ONLY as the
 * test checker and CPU baseline.  Nothing in the shipped library links it.
 *
 * Every function cites the reference file:line it restates.  Layout
 * conventions are the reference's:
 *   - a 2-D Fourier slice of a PxP real image is (P+2) floats x P rows,
 *     interleaved re/im, rows in wrap order (y<0 at P+y)
 *     (libEM/emdata_transform.cpp:118-163);
 *   - the 3-D volume is (P+2) x P x P floats (x fastest) and the weight
 *     ("norm") volume is (P/2+1) x P x P floats
 *     (libEM/reconstructor.cpp:751-802).
 */
#ifndef FOURIER_ORACLE_H
#define FOURIER_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MODE_GAUSS2 0
#define OR_MODE_TRILINEAR 1
#define OR_MODE_GAUSSVAR 2
#define OR_MODE_GAUSS5 3
#define OR_MODE_NEAREST 4
#define OR_MODE_GAUSS3 5
#define OR_MODE_KB 6
#define OR_MODE_GRID5 7
#define OR_MODE_GRID7 8
#define OR_MODE_KBD 9

/* ---- Util helpers (libEM/util.cpp:742-800, libEM/util.h:465-471,816-819) */
void  or_fast_exp_table(float *out1000);
float or_fast_exp(float f);
int   or_hypot_fast_int(int x, int y);
float or_hypot_fast(int x, int y);
int   or_round_f(float x);

/* ---- Transform (libEM/transform.cpp:527-710,1154-1330,1384-1395) -------- */
/* 3x4 row-major float matrices, as Transform::matrix. */
void or_euler_eman(double az, double alt, double phi, float m[12]);
void or_transform_mult(const float a[12], const float b[12], float out[12]);
void or_rotation_only(float m[12]);
void or_vec_times_transform(float x, float y, float z, const float m[12], float out[3]);

/* ---- Symmetry3D (libEM/symmetry.cpp:75-142,1224-1246,1407-1470) -------- */
int or_sym_nsym(const char *sym);
int or_sym_get(const char *sym, float *out /* nsym*12 */);

/* ---- caller prep (programs/e2make3dpar.py:630-654) ----------------------- */
float or_edge_mean(const float *img, int nx, int ny);
void  or_clip_2d(const float *in, int nx, int ny, int x0, int y0, int onx, int ony, float *out);

/* ---- FFT conventions (libEM/emdata_transform.cpp:118-163,392-425) ------- */
void or_fft_r2c_2d(const float *in, int nx, int ny, float *out);
void or_fft_c2r_2d(const float *in, int nx, int ny, float *out);
void or_fft_r2c_1d(const float *in, int nx, float *out);
void or_fft_r2c_3d(const float *in, int nx, int ny, int nz, float *out);
void or_fft_c2r_3d(const float *in, int nx, int ny, int nz, float *out);

/* ---- phase origin (libEM/processor.cpp:6730-6960) ----------------------- */
void or_phaseorigin_tocorner(float *img, int nx, int ny, int nz);
void or_phaseorigin_tocenter(float *img, int nx, int ny, int nz);

/* ---- CTF (libEM/ctf.cpp:964-1090, libEM/ctf.h:291-296, :671-675) -------- */
typedef struct {
    float defocus;   /* um, underfocus positive */
    float bfactor;   /* A^2 */
    float voltage;   /* kV */
    float cs;        /* mm */
    float ampcont;   /* percent */
    float apix;      /* A/pixel */
} or_ctf;
void or_ctf_constants(const or_ctf *c, int ny, float *g1, float *g2, float *acac, float *ds);
void or_ctf_sign_2d(const or_ctf *c, int ny, float *out /* (ny+2) x ny interleaved */);
void or_ctf_amp_2d(const or_ctf *c, int ny, float *out /* (ny+2) x ny interleaved */);

/* ---- preprocess (libEM/reconstructor.cpp:1022-1061) --------------------- */
void or_preprocess_slice(const float *img, int P, float *out /* (P+2)*P */);
void or_prep_particle(const float *raw, int B, int P, const or_ctf *ctf_or_null,
                      float *out /* (P+2)*P */);
void or_prep_particle_xform(const float *raw, int B, int P, const or_ctf *ctf, const float xf[12], float *out);
void or_prep_particle_dfdf(const float *raw, int B, int P, const or_ctf *ctf, float dfdf, const float *xf_or_null,
                           float *out);
int or_xform2d(const float *img, int nx, int ny, const float t[12], float *out);

/* ---- fourier_iter (libEM/reconstructor.cpp:428-682) --------------------- */
void or_iter_insert(int P, float *data, float *norm, const float *slice, const float rot[12], const float *syms,
                    int nsym, float weight, const float *ref_or_null);
void or_iter_ref_from_map(int P, const float *map, float *ref);

/* ---- gauss_fft projector (libEM/projector.cpp:69-295), modes 0-2 -------- */
int or_gauss_fft_project(int P, const float *map, const float t[12], int mode, int returnfft, float *out);

/* ---- volume insertion (libEM/reconstructor.cpp:1122-1210,
 *      libEM/reconstructor_tools.cpp:128-239, libEM/emdata_core.h:368-413) */
void or_insert_slice_work(int P, float *data, float *norm, const float *slice,
                          const float rot[12], const float *syms, int nsym,
                          float weight, int mode, int corners);
int or_insert_slice_ssnr(int P, float *data, float *norm, const float *slice, const float rot[12],
                         const float *syms, int nsym, const float *ssnr, int nssnr, int mode);
void or_insert_slice_work_f64(int P, double *data, double *norm, const float *slice,
                              const float rot[12], const float *syms, int nsym,
                              float weight, int mode, int corners);

/* multi-threaded CPU baseline: T private volumes then summed
 * (programs/e2spa_make3d.py:175-192 merge pattern).  Returns seconds. */
double or_insert_batch_mt(int P, float *data, float *norm, const float *slices,
                          const float *rots, const float *weights, int n,
                          const float *syms, int nsym, int mode, int nthreads);

/* ---- finish (libEM/reconstructor.cpp:315-427,1552-1641) ----------------- */
void or_normalize_threed(int P, float *data, float *norm, int sqrtnorm);
void or_finish(int P, float *data, float *norm, int sqrtnorm, int doift,
               float *out_real /* P^3, only if doift */);

/* ---- slice agreement (libEM/reconstructor.cpp:1212-1381,1448-1461) ------ */
void or_determine_slice_agreement(int P, float *data, float *norm, float *slice, int mode,
                                  const float rot[12], const float *syms, int nsym,
                                  float weight, int sub, float apix, float out[5]);

/* ---- seed (libEM/reconstructor.cpp:820-918) ------------------------------ */
void or_setup_seed(int P, float *data, float *norm, const float *seed, float seed_weight);

#ifdef __cplusplus
}
#endif
void or_projection(int P, const float *data, const float *norm, const float euler[12], int ret_fourier, float *out);

double or_bessel_i0_scaled(double x);

#endif

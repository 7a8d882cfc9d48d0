/*
 * fourier_hip.h -- C ABI of the MI355X direct Fourier inversion
 * reconstructor ("fourier_hip").
 *
 * The entry points are the calls libEM's Reconstructor surface makes on the
 * "fourier" Reconstructor (libEM/reconstructor.h:108-206,
 * libEM/reconstructor.cpp:686-1641), flattened to plain pointers and sizes
 * so a libEM-side `FourierHipReconstructor : Reconstructor`, a Boost.Python /
 * pybind11 module or ctypes can bind them.  No C++ exceptions cross this
 * boundary: every call returns an int status (FH_OK = 0), and the message of
 * the last failure on a handle is fh_last_error(h).
 *
 * Array layouts are the reference's:
 *   image   : P x P float32, row-major (y, x)
 *   slice   : P rows x (P+2) floats, interleaved re/im, rows in wrap order
 *             (libEM/emdata_transform.cpp:118-163 do_fft_inplace layout)
 *   volume  : P x P x (P+2) floats [z][y][x re/im]  (reconstructor.cpp:751-787)
 *   weights : P x P x (P/2+1) floats                  (reconstructor.cpp:798-802)
 *   xform   : 12 floats, the 3x4 row-major Transform::matrix
 *             (libEM/transform.h:503)
 * where P is the padded box ("size" parameter).
 *
 * Thread safety: one handle may be called from several host threads at
 * once (e2make3dpar.py:355-381 does this); calls on a handle are
 * serialised internally.  Different handles run on their own HIP streams
 * and overlap on the device.
 */
#ifndef FOURIER_HIP_H
#define FOURIER_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the reference's exception types (libEM/exception.h:67-380) */
#define FH_OK 0
#define FH_SKIPPED (-1)            /* insert_slice with weight 0 (reconstructor.cpp:1088) */
#define FH_ERR_INVALID_PARAM 1     /* InvalidParameterException */
#define FH_ERR_NOT_EXISTING 2      /* NotExistingObjectException (unknown sym/mode) */
#define FH_ERR_IMAGE_DIM 3         /* ImageDimensionException */
#define FH_ERR_NULL 4              /* NullPointerException */
#define FH_ERR_DEVICE 5            /* HIP runtime error */
#define FH_ERR_STATE 6             /* call before setup() or after finish() */
#define FH_ERR_NOMEM 7             /* device allocation refused */

/* memory kinds for batch inputs */
#define FH_HOST 0
#define FH_DEVICE 1

typedef struct fh_reconstructor fh_reconstructor;

/* Construction parameters: the "fourier" Reconstructor's TypeDict
 * (libEM/reconstructor.h:501-517) restricted to what this path implements. */
typedef struct {
    int nx, ny, nz;        /* "size": padded real-space box, nx==ny==nz, even */
    const char *sym;       /* "sym": c<n>, d<n>, h<n>[:1:daz:tz], tet, oct, icos (symmetry.cpp:75-142) */
    const char *mode;      /* "mode": gauss_2, trilinear, gauss_var, gauss_5, gauss_3, nearest_neighbor,
                              kaiser_bessel, kaiser_bessel_derived, gridding_5, gridding_7
                              (reconstructor_tools.cpp:42-53) */
    int corners;           /* "corners" (reconstructor.cpp:1175-1181) */
    int sqrtnorm;          /* "sqrtnorm" (reconstructor.cpp:1564) */
    int usessnr;           /* "usessnr": weight slices by their class_ssnr (not with corners) */
    int verbose;           /* "verbose" */
    int device;            /* HIP device ordinal */
} fh_params;

/* CTF constants shared by a batch; defocus comes per particle.  Used for the
 * CTF_SIGN phase flip of e2make3dpar.py:704-714 (ctf.cpp:1076-1090): the
 * padded particle's FFT times CTF_SIGN at its defocus, back to real space,
 * before preprocess_slice's 2-D resampling.  With dfdf (n values, host) the
 * flip is the product of the signs at defocus and at defocus + dfdf, the
 * defocus-offset flip of e2make3dpar.py:698-709 and e2spa_make3d.py:316-327
 * (where dfdf is 0 the two signs cancel: no flip). */
typedef struct {
    float voltage;       /* kV */
    float cs;            /* mm */
    float ampcont;       /* % amplitude contrast */
    float apix;          /* A/pixel */
    const float *dfdf;   /* optional per-particle defocus offsets (um), or NULL */
} fh_ctf;

/* Tuning knobs of the batched insertion (picked by the autotuner).  The
 * stencil origins of the unfolded frequency grid are cut into 8^3 bricks; a
 * work item is one brick and up to slices_per_item slices sharing a dominant
 * normal axis. */
typedef struct {
    int waves_per_block; /* independent item workers per workgroup: 1 to 4 */
    int slices_per_item; /* slices a work item adds before writing its partial brick */
    int chunk;           /* particles per batched launch; 0 = as many as fit in 48 GiB of staging */
    int blocks_per_cu;   /* persistent grid: workgroups per compute unit */
    int graph;           /* 1: device-resident batches run each chunk as one captured hipGraph */
} fh_tuning;

/* Timings and counters of the last batched call (device time, ms). */
typedef struct {
    double prep_ms, bin_ms, insert_ms, total_ms;
    double insert_kernel_ms;     /* the brick scatter kernel alone */
    double merge_ms;             /* the partial-brick merge into the volume */
    long long slices, pairs, items, chunks;
} fh_stats;

/* Reconstructors.get("fourier_hip", {...}) -- Factory<Reconstructor>::get
 * (libEM/emobject.h:801-834): validates the parameters. */
int fh_create(const fh_params *p, fh_reconstructor **out);
int fh_destroy(fh_reconstructor *h);
/* message of the last failure on h (h == NULL: last failure of fh_create
 * on this thread) */
const char *fh_last_error(const fh_reconstructor *h);

/* FourierReconstructor::setup (reconstructor.cpp:730-818) */
int fh_setup(fh_reconstructor *h);
/* FourierReconstructor::setup_seed (reconstructor.cpp:820-918); seed is a
 * host volume in the reference layout */
int fh_setup_seed(fh_reconstructor *h, const float *seed, float seed_weight);
/* FourierReconstructor::setup_seedandweights (reconstructor.cpp:920-996) */
int fh_setup_seedandweights(fh_reconstructor *h, const float *seed, const float *weights);
/* FourierReconstructor::clear (reconstructor.cpp:999-1020) */
int fh_clear(fh_reconstructor *h);

/* FourierReconstructor::preprocess_slice (reconstructor.cpp:1022-1061) for
 * a pure rotation: host P x P image -> host slice. */
int fh_preprocess_slice(fh_reconstructor *h, const float *img, float *slice_out);
/* The same with the Transform preprocess_slice receives: its 2-D part
 * (translation, scale, x mirror; reconstructor.cpp:1031-1035) resamples the
 * image first ("xform" processor, processor.cpp:11025-11075). */
int fh_preprocess_slice_xform(fh_reconstructor *h, const float *img, const float xform[12], float *slice_out);

/* FourierReconstructor::insert_slice (reconstructor.cpp:1064-1119).
 * data is a host P x P image (is_preprocessed = 0) or a host slice
 * (is_preprocessed = 1, the "reconstruct_preproc" attribute).
 * Returns FH_SKIPPED for weight 0. */
int fh_insert_slice(fh_reconstructor *h, const float *data, int is_preprocessed,
                    const float xform[12], float weight);

/* insert_slice of a slice carrying a "class_ssnr" attribute
 * (reconstructor.cpp:1080-1088,1158-1184): with usessnr set, each pixel's
 * weight is max(0, class_ssnr[int(r * 2 (len-1) / P)]) at slice radius
 * r = hypot_fast_int(x, y) and `weight` is ignored; with usessnr off the
 * class_ssnr is ignored and `weight` applies, as in the reference.  (Without
 * class_ssnr, fh_insert_slice on a usessnr handle returns FH_SKIPPED.) */
int fh_insert_slice_ssnr(fh_reconstructor *h, const float *data, int is_preprocessed,
                         const float xform[12], float weight, const float *class_ssnr, int ssnr_len);

/* Batched insertion of n raw particles (box x box, box <= P) with the
 * caller prep of e2make3dpar.py:630-654,704-716 fused in: edge-mean
 * subtract, centre zero-pad to P, optional CTF_SIGN flip (defocus != NULL
 * and ctf != NULL), phase-origin shift, r2c FFT, scale, then gridding of all
 * n x nsym slices.  imgs/xforms/weights/defocus live in `where` memory. */
int fh_insert_batch(fh_reconstructor *h, const float *imgs, int n, int box,
                    const float *xforms, const float *weights, const float *defocus,
                    const fh_ctf *ctf, int where);

/* fh_insert_batch for particles that all carry a class_ssnr of ssnr_len
 * values (host, n x ssnr_len): the per-radius weighting of
 * fh_insert_slice_ssnr on a usessnr handle, `weights` otherwise.  (On a
 * usessnr handle fh_insert_batch and fh_insert_plain return FH_SKIPPED:
 * particles without class_ssnr have weight 0.) */
int fh_insert_batch_ssnr(fh_reconstructor *h, const float *imgs, int n, int box,
                         const float *xforms, const float *weights, const float *class_ssnr,
                         int ssnr_len, const float *defocus, const fh_ctf *ctf, int where);

/* The prep stage of fh_insert_batch alone: n raw particles -> n slices
 * (host, n x P x (P+2) floats).  The caller prep of e2make3dpar.py:630-654,
 * 704-714 followed by preprocess_slice (reconstructor.cpp:1022-1061). */
int fh_prep_batch(fh_reconstructor *h, const float *imgs, int n, int box, const float *defocus,
                  const fh_ctf *ctf, float *slices_out, int where);

/* Same contract, structured like the reference's CUDA path
 * (reconstructor.cpp:1074-1078,1141-1157; cuda/cuda_reconstructor.cu:171-193):
 * per particle a host->device copy, one insertion launch per symmetry op,
 * a synchronise after every launch.  The "plain" GPU form. */
int fh_insert_plain(fh_reconstructor *h, const float *imgs, int n, int box,
                    const float *xforms, const float *weights, const float *defocus,
                    const fh_ctf *ctf);
/* The plain form with the particles where `where` says: FH_DEVICE skips the
 * per-particle copy (launches and synchronisations only). */
int fh_insert_plain_where(fh_reconstructor *h, const float *imgs, int n, int box,
                          const float *xforms, const float *weights, const float *defocus,
                          const fh_ctf *ctf, int where);

/* FourierReconstructor::determine_slice_agreement (reconstructor.cpp:1212-1257)
 * for one slice: out = {reconstruct_norm, reconstruct_absqual,
 * reconstruct_absqual_lowres, reconstruct_qual, reconstruct_weight}. */
int fh_determine_slice_agreement(fh_reconstructor *h, const float *data, int is_preprocessed,
                                 const float xform[12], float weight, int sub, float apix,
                                 float out[5]);

/* determine_slice_agreement in the reference's own order for a slice that
 * carries class_ssnr (reconstructor.cpp:1212-1257): with sub, the first
 * do_insert_slice_work(-weight) takes its weight < 0 SSNR branch
 * (reconstructor.cpp:1160-1184) and ADDS the slice with the per-radius
 * weights max(0, class_ssnr[int(r * 2 (len-1) / P)]) instead of removing it,
 * the compare follows, and the slice is added again with +weight: the
 * volume keeps both copies.  sub without class_ssnr fails with
 * FH_ERR_NOT_EXISTING, the reference's NotExistingObjectException.
 * fh_determine_slice_agreement (sub) removes the slice's own contribution
 * instead -- what the reference code intends and what the refine flow on
 * raw particles needs.  sub = 0 is fh_determine_slice_agreement. */
int fh_determine_slice_agreement_ssnr(fh_reconstructor *h, const float *data, int is_preprocessed,
                                      const float xform[12], float weight, int sub, float apix,
                                      const float *class_ssnr, int ssnr_len, float out[5]);

/* Batched agreement of n raw particles against the current volume, each
 * compared as if its own contribution were removed (sub) -- the
 * e2make3d.py:447-491 reweighting pass.  out is n x 5 (host or device per
 * `where`). */
int fh_agreement_batch(fh_reconstructor *h, const float *imgs, int n, int box,
                       const float *xforms, const float *weights, const float *defocus,
                       const fh_ctf *ctf, int sub, float apix, float *out, int where);

/* No counterpart in the reference: determine_slice_agreement for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same outputs as
 * fh_determine_slice_agreement on an already preprocessed slice, which the
 * caller passes in device memory (slice_dev: P x (P+2) floats); out (host)
 * as there.  It calls no allocator: its working memory (rotations, the
 * compare kernel's partial sums) is a fixed 2 MiB + 4 KiB device array per
 * GPU, so it refuses (FH_ERR_INVALID_PARAM) sizes and symmetries needing
 * more than 2 MiB of partial sums (8 doubles per 256 compare pixels and
 * symmetry operator) or more than 64 operators.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int fh_determine_slice_agreement_noalloc(fh_reconstructor *h, const float *slice_dev,
                                         const float xform[12], float weight, int sub,
                                         float apix, float out[5]);

/* fourier_iter (FourierIterReconstructor, reconstructor.cpp:428-682) is the
 * handle created with mode "fourier_iter": its own 6x6x6 Gaussian insertion;
 * fh_setup_seed / fh_setup_seedandweights keep a complex seed (P x P x (P+2))
 * as its reference volume (weights ignored), which moves every inserted value
 * by the reference's local difference; determine_slice_agreement returns
 * without comparing; finish() as fourier's.  fh_setup_seed_map sets the
 * reference from a real P^3 map (host): do_fft with the phase origin moved
 * to the corner, on the GPU (e2make3dpar.py:370-372 seeds each iteration
 * with the previous map). */
int fh_setup_seed_map(fh_reconstructor *h, const float *map);

/* The gauss_fft projector (GaussFFTProjector::project3d, projector.cpp:
 * 69-295; e2make3dpar.py:716-721 projects the reference with it), modes 0
 * (trilinear, the default), 1 (nearest) and 2 (Gaussian): fh_gauss_fft_volume
 * keeps the transform of a real P^3 map (host) on the handle's device;
 * fh_gauss_fft_project makes n projections at the Transforms xforms (n x 12,
 * host): returnfft -> n x P x (P+2) Fourier slices with the translation's
 * phase ramp, else n x P x P images (inverse FFT, phase origin centred, the
 * translation resampled, math.gausskernelfix).  Mirrored Transforms are
 * refused. */
int fh_gauss_fft_volume(fh_reconstructor *h, const float *map);
int fh_gauss_fft_project(fh_reconstructor *h, int n, const float *xforms, int mode, int returnfft, float *out);

/* The task merge of e2spa_make3d.py:175-192: after fh_setup(h), each task's
 * finish(doift=0) volume `threed` (P x P x (P+2) floats) times its normout
 * (P x P x (P/2+1), fh_finish's norm_out) is added to h's volume and the
 * normout to h's weights (host or device arrays per `where`); then
 * fh_merge_finish multiplies the sum by the reciprocal of the summed weights
 * (0 where 0), runs do_ift_inplace, depad and the phase origin to the centre
 * into out (host, P^3 floats) and releases the volume, as fh_finish does. */
int fh_merge_add(fh_reconstructor *h, const float *threed, const float *normv, int where);
int fh_merge_finish(fh_reconstructor *h, float *out);

/* FourierReconstructor::projection (reconstructor.cpp:1383-1445): the
 * Fourier slice of the accumulating volume at the rotation of xform (value
 * over weight of the nearest voxel, weights below 0.005 left out), shifted
 * by xform's translation.  out (host): ret_fourier -> P x (P+2) slice floats,
 * else the P x P real-space image (phase origin to centre, inverse FFT). */
int fh_projection(fh_reconstructor *h, const float xform[12], int ret_fourier, float *out);

/* FourierReconstructor::finish (reconstructor.cpp:1552-1641):
 * normalize_threed, then (doift) 3-D c2r, x 1/P^3, depad, phase origin to
 * centre.  out (host): P^3 floats if doift, else the normalised complex
 * volume.  norm_out (optional, host): the fixed-up weights ("normout").
 * The volume is released afterwards, as the reference does. */
int fh_finish(fh_reconstructor *h, int doift, float *out, float *norm_out);

/* Copy the accumulating (complex, weight) state to host buffers
 * (the normout/savenorm export of reconstructor.cpp:1621-1631 plus the
 * complex sums, for checkpointing with fh_setup_seedandweights). */
int fh_get_state(fh_reconstructor *h, float *data, float *weights);

int fh_set_tuning(fh_reconstructor *h, const fh_tuning *t);
int fh_get_tuning(const fh_reconstructor *h, fh_tuning *t);
int fh_get_stats(const fh_reconstructor *h, fh_stats *s);
/* wait for all queued device work of this handle */
int fh_sync(fh_reconstructor *h);
/* Symmetry3D::get_symmetries (libEM/symmetry.cpp:1240-1246): the rotation
 * operators of a symmetry name as row-major 3x3 matrices; returns their
 * count (out9 NULL: count only), -FH_ERR_NOT_EXISTING for an unknown name,
 * -FH_ERR_INVALID_PARAM when max_ops is too small.  No device needed. */
int fh_get_symmetries(const char *sym, float *out9, int max_ops);
/* device bytes this handle holds / would hold after setup for size P
 * (chunk 0: the automatic chunk for particles of box P) */
long long fh_footprint_bytes(int P, int nsym, int chunk);

/* ------------------------------------------------------------------------
 * Serving: a request loop that accepts reconstruction jobs of mixed box
 * sizes as they arrive, admits them by HBM footprint, keeps each job's
 * volume resident on the device, and interleaves the jobs' chunks on the
 * GPU (continuous batching: a job admitted while others run starts
 * immediately, its work overlapping theirs on separate streams).  No
 * counterpart in the reference, whose task queue (libpyEM/EMAN2PAR.py)
 * runs each make3d job in its own process; this is that queue for one GPU.
 * ---------------------------------------------------------------------- */
typedef struct fh_server fh_server;

/* A job: n particles (box x box, in `where` memory; must stay valid until
 * the job completes), host xforms (n x 12), weights (n), optional defocus
 * (n) + ctf, the reconstructor parameters, and where to put the result. */
typedef struct {
    fh_params params;
    const float *imgs;
    int n, box, where;
    const float *xforms, *weights, *defocus;
    fh_ctf ctf;
    int use_ctf;
    float *out_map;    /* host P^3 floats, filled by finish(doift=1) */
} fh_job;

#define FH_JOB_QUEUED 0
#define FH_JOB_RUNNING 1
#define FH_JOB_DONE 2
#define FH_JOB_FAILED 3
#define FH_JOB_REJECTED 4
#define FH_JOB_CANCELLED 5

typedef struct {
    int state;
    int error;                  /* status code when FAILED / REJECTED */
    double t_submit, t_admit, t_done;  /* seconds on the server's clock */
    long long particles_done;
    long long footprint_bytes;
} fh_job_status;

/* hbm_budget_bytes: device memory the server may commit to admitted jobs;
 * max_active: jobs running at once (each has its own stream and worker). */
int fh_server_create(int device, long long hbm_budget_bytes, int max_active, int chunk, fh_server **out);
int fh_server_submit(fh_server *s, const fh_job *job, int *job_id);
int fh_server_poll(fh_server *s, int job_id, fh_job_status *st);
int fh_server_cancel(fh_server *s, int job_id);
/* block until every submitted job has left the queue and finished */
int fh_server_drain(fh_server *s);
double fh_server_clock(fh_server *s);
int fh_server_destroy(fh_server *s);
const char *fh_server_last_error(const fh_server *s);

#ifdef __cplusplus
}
#endif
#endif

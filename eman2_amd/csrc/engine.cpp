// fourier_hip host engine: the "fourier" Reconstructor surface
// (libEM/reconstructor.h:108-206, libEM/reconstructor.cpp:686-1641) over the
// HIP kernels, exported through the C ABI of include/fourier_hip.h.
//
// The handle owns the device-resident volume (complex sums + weights in the
// reference layout), the particle staging buffers of one batched chunk and
// the binning buffers.  Every call on a handle is serialised by its mutex and
// issued on the handle's own HIP stream (default priority).
#include "fourier_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fh_launch.hpp"

using namespace fh;

namespace {

// ------------------------------------------------------------------ math
// Util::round / fast_floor / apply_precision (libEM/util.h:465-471,874-880,1309-1314)
int fast_floor(float x) { return x < 0 ? ((int)x - 1) : (int)x; }
void apply_precision(float &value, float precision)
{
    float c = ceilf(value);
    float f = (float)fast_floor(value);
    if (fabsf(value - c) < precision) value = c;
    else if (fabsf(value - f) < precision) value = f;
}
const float ERR_LIMIT = 0.000001f;  // Transform::ERR_LIMIT (libEM/transform.cpp:57)

// Transform::get_determinant (libEM/transform.cpp:1310-1322)
float determinant(const float *m)
{
    double det2;
    det2 = m[0] * ((double)m[5] * m[10] - (double)m[9] * m[6]);
    det2 -= m[1] * ((double)m[4] * m[10] - (double)m[8] * m[6]);
    det2 += m[2] * ((double)m[4] * m[9] - (double)m[8] * m[5]);
    float det = (float)det2;
    apply_precision(det, ERR_LIMIT);
    return det;
}

// Transform::get_scale (libEM/transform.cpp:1176-1187)
float get_scale(const float *m)
{
    float d = determinant(m);
    if (d < 0) d *= -1;
    float scale = powf(d, 1.0f / 3.0f);
    int is = (int)scale;
    float res = scale - (float)is;
    if (res < ERR_LIMIT) scale = (float)is;
    apply_precision(scale, ERR_LIMIT);
    return scale;
}

// set_scale(1); set_mirror(false); set_trans(0,0,0) as insert_slice does
// (libEM/reconstructor.cpp:1104-1106); returns the 3x3 rotation.
void rotation_only(const float *in12, float *out9)
{
    float m[12];
    memcpy(m, in12, sizeof(m));
    float old_scale = get_scale(m);
    float n_scale = 1.0f;
    apply_precision(n_scale, ERR_LIMIT);
    float corrected = n_scale / old_scale;
    if (corrected != 1.0)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) m[i * 4 + j] *= corrected;
    if (determinant(m) < 0)
        for (int j = 0; j < 4; j++) m[j] *= -1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out9[3 * i + j] = m[4 * i + j];
}

// eman Euler -> 3x3 (libEM/transform.cpp:650-667)
void euler_eman(double az, double alt, double phi, float *m9)
{
    const double d2r = 3.141592653589793238462643383279502884197169399 / 180.0;
    double azp = az * d2r, altp = alt * d2r, phip = phi * d2r;
    m9[0] = (float)(cos(phip) * cos(azp) - cos(altp) * sin(azp) * sin(phip));
    m9[1] = (float)(cos(phip) * sin(azp) + cos(altp) * cos(azp) * sin(phip));
    m9[2] = (float)(sin(altp) * sin(phip));
    m9[3] = (float)(-sin(phip) * cos(azp) - cos(altp) * sin(azp) * cos(phip));
    m9[4] = (float)(-sin(phip) * sin(azp) + cos(altp) * cos(azp) * cos(phip));
    m9[5] = (float)(sin(altp) * cos(phip));
    m9[6] = (float)(sin(altp) * sin(azp));
    m9[7] = (float)(-sin(altp) * cos(azp));
    m9[8] = (float)cos(altp);
}

// Symmetry3D::get_symmetries (libEM/symmetry.cpp:75-142 name parse,
// :1240-1246): c<n> (CSym::get_sym :1407-1418), d<n> (DSym::get_sym
// :1452-1470), h<n>[:<nstart>:<daz>:<tz>[:<maxtilt>]] (HSym::get_sym
// :1648-1671, nstart 1 only, as the reference), tet / oct / icos / icos5
// (TetrahedralSym :2007-2030, OctahedralSym :1950-1975, IcosahedralSym
// :1822-1866).  Each operator is the 3x3 rotation of the eman-Euler
// Transform the reference builds; the helical z translation is dropped
// because insertion uses only the 3x3 part (transform.h:547-553).
bool sym_ops(const std::string &name, std::vector<float> &out)
{
    if (name.empty() || name.size() > 63) return false;
    std::string low(name);
    for (char &ch : low) ch = (char)tolower((unsigned char)ch);
    out.clear();
    auto push = [&](double az, double alt, double phi) {
        float m[9];
        euler_eman(az, alt, phi, m);
        out.insert(out.end(), m, m + 9);
    };
    const double rad2deg = 180.0 / 3.141592653589793238462643383279502884197169399;  // EMConsts::rad2deg
    if (low == "tet") {
        // faces: one along z (az 0, phi 0/120/240), three at acos(-1/3)
        // (az 0/120/240, phi 60/180/300); ::acos(double) of the float -1/3
        const float l1 = (float)(rad2deg * acos((double)(-1.0f / 3.0f)));
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 3; c++) {
                if (r == 0) push(0.0, 0.0, 120.0 * c);
                else push(120.0 * (r - 1), (double)l1, 60.0 + 120.0 * c);
            }
        return true;
    }
    if (low == "oct") {
        // a cube face along z: alt 0, four alt-90 rows (az 0/90/180/270), alt 180;
        // phi 0/90/180/270 in each row
        for (int r = 0; r < 6; r++)
            for (int c = 0; c < 4; c++) {
                double az = (r >= 1 && r <= 4) ? 90.0 * (r - 1) : 0.0;
                double alt = r == 0 ? 0.0 : (r == 5 ? 180.0 : 90.0);
                push(az, alt, 90.0 * c);
            }
        return true;
    }
    if (low == "icos" || low == "icos5") {
        // a pentagon normal to z; levels 0, atan(2), 180 - atan(2), 180 (::atan(double)
        // of the float 2); the 234-degree alignment offset adds 90 to az and
        // subtracts 90 from phi (symmetry.cpp:1820, 1845-1849)
        const float l1 = (float)(rad2deg * atan((double)2.0f));
        const float l2 = 180.0f - l1;
        for (int r = 0; r < 12; r++)
            for (int c = 0; c < 5; c++) {
                float ring = c == 0 ? 0.0f : 360.0f - 72.0f * (float)c;  // 0, 288, 216, 144, 72
                float az, alt, phi;
                if (r == 0) { az = 0.0f; alt = 0.0f; phi = ring; }
                else if (r <= 5) { az = 72.0f * (float)(r - 1); alt = l1; phi = c == 0 ? 36.0f : 396.0f - 72.0f * (float)c; }
                else if (r <= 10) { az = 36.0f + 72.0f * (float)(r - 6); alt = l2; phi = ring; }
                else { az = 0.0f; alt = 180.0f; phi = ring; }
                push((double)(az + 90.0f), (double)alt, (double)(phi - 90.0f));
            }
        return true;
    }
    const char k = low[0];
    if (k == 'h') {
        // the reference's two sscanf forms and defaults (nsym 1, nstart 1, daz 0, tz 0, maxtilt 90)
        int nsym = 1, nstart = 1;
        float daz = 0.0f, tz = 0.0f, maxtilt = 90.0f;
        if (sscanf(low.c_str() + 1, "%d:%d:%f:%f:%f", &nsym, &nstart, &daz, &tz, &maxtilt) < 4)
            sscanf(low.c_str() + 1, "%d,%d,%f,%f,%f", &nsym, &nstart, &daz, &tz, &maxtilt);
        if (nstart != 1 || nsym <= 0) return false;
        for (int i = 0; i < nsym; i++) {
            int ii = (i + 1) / 2;  // 0, 1, -1, 2, -2, ...
            if (i > 1 && i % 2 == 0) ii = -ii;
            double az = (double)(ii % nstart) * (360.0 / nstart) + floor((double)((float)ii / (float)nstart)) * (double)daz;
            push(az, 0.0, 0.0);
        }
        return true;
    }
    if (k != 'c' && k != 'd') return false;
    if (low.size() < 2) return false;
    for (size_t i = 1; i < low.size(); i++)
        if (!isdigit((unsigned char)low[i])) return false;
    int n = atoi(low.c_str() + 1);
    if (n <= 0) return false;
    if (k == 'c') {
        for (int i = 0; i < n; i++) {
            float az = (float)(i % n) * 360.0f / (float)n;
            push((double)az, 0.0, 0.0);
        }
        return true;
    }
    int nsym = 2 * n;
    for (int i = 0; i < nsym; i++) {
        float az, alt;
        if (i >= nsym / 2) {
            az = (float)((i % nsym) - nsym / 2) * 360.0f / (float)(nsym / 2);
            alt = 180.0f;
        } else {
            az = (float)(i % nsym) * 360.0f / (float)(nsym / 2);
            alt = 0.0f;
        }
        push((double)az, (double)alt, 0.0);
    }
    return true;
}

// preprocess_slice's 2-D transform (libEM/reconstructor.cpp:1031-1035):
// tmp = t with its rotation reset by set_rotation({"type": "eman"}), which
// keeps scale and x mirror (transform.cpp:527-707: scale and mirror taken
// first, the eman (0,0,0) matrix, scale re-applied, row 0 negated for a
// mirror) and the translation.  is_identity(tmp) (transform.cpp:222-235)
// -> the image is used as is; otherwise the "xform" processor samples it at
// inverse(tmp) . (x, y) (Transform::invert, transform.cpp:1324-1356, in
// double).  Returns false for identity, else inverse rows m00 m01 m03 m10
// m11 m13.
bool xform2d_inverse(const float *m, float inv6[6])
{
    // fast path: no translation beyond ERR_LIMIT and determinant 1 (a pure
    // rotation) make tmp the identity
    if (fabsf(m[3]) < ERR_LIMIT && fabsf(m[7]) < ERR_LIMIT && fabsf(m[11]) < ERR_LIMIT && determinant(m) == 1.0f)
        return false;
    const float det0 = determinant(m);
    const bool mirror = det0 < 0;
    const float scale = get_scale(m);
    float t[12];
    memcpy(t, m, sizeof(t));
    // the eman (0, 0, 0) rotation, evaluated once (the same trig as set_rotation)
    static const std::vector<float> e0 = [] {
        std::vector<float> e(9);
        euler_eman(0.0, 0.0, 0.0, e.data());
        return e;
    }();
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[4 * i + j] = e0[3 * i + j];
    if (scale != 1.0f)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) t[4 * i + j] *= scale;
    if (mirror)
        for (int j = 0; j < 3; j++) t[j] *= -1.0f;
    bool ident = true;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) {
            float c = t[4 * i + j];
            apply_precision(c, ERR_LIMIT);
            if (i == j ? c != 1.0f : c != 0.0f) ident = false;
        }
    if (ident) return false;
    const double m00 = t[0], m01 = t[1], m02 = t[2], m10 = t[4], m11 = t[5], m12 = t[6];
    const double m20 = t[8], m21 = t[9], m22 = t[10], v0 = t[3], v1 = t[7], v2 = t[11];
    const double cof00 = m11 * m22 - m12 * m21, cof11 = m22 * m00 - m20 * m02, cof01 = m10 * m22 - m20 * m12;
    const double cof02 = m10 * m21 - m20 * m11, cof10 = m01 * m22 - m02 * m21, cof20 = m01 * m12 - m02 * m11;
    const double cof21 = m00 * m12 - m10 * m02;
    const double det = m00 * cof00 + m02 * cof02 - m01 * cof01;
    inv6[0] = (float)(cof00 / det);
    inv6[1] = -(float)(cof10 / det);
    inv6[2] = (float)((-cof00 * v0 + cof10 * v1 - cof20 * v2) / det);
    inv6[3] = -(float)(cof01 / det);
    inv6[4] = (float)(cof11 / det);
    inv6[5] = (float)((cof01 * v0 - cof11 * v1 + cof21 * v2) / det);
    return true;
}

// t3d = a * b in FP32, reference order (libEM/transform.cpp:1384-1395)
void compose9(const float *a, const float *b, float *r)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[3 * i + j] = a[3 * i + 0] * b[0 * 3 + j] + a[3 * i + 1] * b[1 * 3 + j] + a[3 * i + 2] * b[2 * 3 + j];
}

// CTF constants (libEM/ctf.h:291-296, libEM/ctf.cpp:671-675, :985-992)
CtfConst ctf_constants(const fh_ctf &c, int ny)
{
    CtfConst k{};
    float lambda = 12.2639f / sqrtf(c.voltage * 1000.0f + 0.97845f * c.voltage * c.voltage);
    float phase;
    if (c.ampcont > -100.0 && c.ampcont <= 100.0) phase = (float)asin(c.ampcont / 100.0);
    else if (c.ampcont > 100.0) phase = (float)(M_PI - asin(2.0f - c.ampcont / 100.0));
    else phase = (float)(-M_PI - asin(-2.0 - c.ampcont / 100.0));
    k.ds = 1.0f / (c.apix * ny);
    k.g1 = (float)(M_PI / 2.0 * c.cs * 1.0e7 * powf(lambda, 3.0f));
    k.g2 = (float)(M_PI * lambda * 10000.0);
    k.acac = (float)(M_PI / 2.0 - phase);
    return k;
}

// sizes whose prime factors the FFT codelets cover: every even entry of
// EMAN2's good_box_sizes (libpyEM/EMAN2.py:660) has factors 2, 3, 5, 7, 11, 13
bool fft_size_ok(int n)
{
    if (n < 2) return false;
    for (int p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

thread_local std::string g_create_error;

}  // namespace

// ====================================================================== handle
struct fh_reconstructor {
    std::mutex mu;
    std::string err;
    // parameters
    int P = 0, H = 0, device = 0;
    std::string sym, mode_name;
    int mode = 0, corners = 0, sqrtnorm = 0, verbose = 0, usessnr = 0;
    std::vector<float> syms;  // nsym x 9
    int nsym = 1;
    // device state
    hipStream_t stream = nullptr;
    bool ready = false;  // setup() called and finish() not yet
    float2 *data = nullptr;
    float *norm = nullptr;
    float2 *ref = nullptr;      // fourier_iter: the reference volume of setup_seed (or null)
    float2 *pvol = nullptr;     // gauss_fft projector: the map's transform (phase origin at the corner)
    float *exptab = nullptr, *rtab = nullptr, *syms_d = nullptr;
    float2 *tw = nullptr;
    FftPlan plan{};
    float scale = 1.f;
    int cus = 256;
    // chunk staging
    int cap_n = 0, cap_B = 0;
    float *imgs = nullptr, *means = nullptr, *rots = nullptr, *wts = nullptr, *dfs = nullptr;
    float *dfs2 = nullptr;      // second defocus (defocus + dfdf) of the two-sign flip
    float2 *slices = nullptr;
    float2 *inter = nullptr;  // column-first prep intermediate of inter_n particles
    int inter_n = 0;
    float *t3d = nullptr, *wq = nullptr;
    float *rw = nullptr;        // class_ssnr radial weights, (P/2+1) per particle (usessnr)
    float *rw1 = nullptr;       // one slice's class_ssnr radial weights (agreement in the reference's order)
    float *ximgs = nullptr;     // 2-D transformed padded particles, P x P each
    float *ximgs2 = nullptr;    // second P x P staging: the flipped images' resampling (CTF with a 2-D part)
    float *xfd = nullptr;       // their inverse 2-D transforms, 6 per particle
    int ximgs_cap = 0, ximgs2_cap = 0;
    double *partial = nullptr;  // agreement partial sums
    double *acc8 = nullptr;     // batched agreement sums, 8 per particle
    int acc8_cap = 0;
    // binning
    BrickGeom geom{};
    BinBuffers bins{};
    long long bins_cap_slices = 0;
    fh_tuning tune{2, 1024, 0, 8, 0};
    fh_stats stats{};
    hipEvent_t ev[6] = {};
    // pinned staging for large device-to-host copies (two buffers, one event each)
    float *pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {};
};

#define FH_TRY(h, expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) {                                                            \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                  \
            return FH_ERR_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

static int fail(fh_reconstructor *h, int code, const std::string &msg)
{
    h->err = msg;
    return code;
}

static size_t vol_elems(int P) { return (size_t)P * P * (P / 2 + 1); }

// Device-to-host copy of `bytes` into caller (pageable) memory: through two
// pinned staging buffers, the DMA of one chunk running while several host
// threads copy the previous chunk out; small copies go directly.
constexpr size_t PIN_CHUNK = 64ull << 20;
static void par_memcpy(char *dst, const char *src, size_t len)
{
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, len >> 23));
    if (nt == 1) {
        memcpy(dst, src, len);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (len + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        const size_t a = t * per, b = std::min(len, a + per);
        if (a < b) th.emplace_back([=] { memcpy(dst + a, src + a, b - a); });
    }
    for (auto &x : th) x.join();
}

// the two staging buffers and their events (recorded once, so that a first
// wait on them returns at once)
static hipError_t ensure_pinned(fh_reconstructor *h)
{
    hipError_t e;
    for (int i = 0; i < 2; i++) {
        if (!h->pin[i] && (e = hipHostMalloc((void **)&h->pin[i], PIN_CHUNK, hipHostMallocDefault)) != hipSuccess)
            return e;
        if (!h->pin_ev[i]) {
            if ((e = hipEventCreateWithFlags(&h->pin_ev[i], hipEventDisableTiming)) != hipSuccess) return e;
            if ((e = hipEventRecord(h->pin_ev[i], h->stream)) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

static hipError_t d2h_staged(fh_reconstructor *h, void *dst, const void *src, size_t bytes)
{
    hipError_t e;
    if (bytes < (16ull << 20)) {
        if ((e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream)) != hipSuccess) return e;
        return hipStreamSynchronize(h->stream);
    }
    if ((e = ensure_pinned(h)) != hipSuccess) return e;
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    const size_t nchunk = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
    for (size_t c = 0; c <= nchunk; c++) {
        if (c < nchunk) {  // the DMA of chunk c into buffer c % 2 (free: chunk c-2 was copied out)
            const size_t off = c * PIN_CHUNK, len = std::min(PIN_CHUNK, bytes - off);
            if ((e = hipMemcpyAsync(h->pin[c & 1], s + off, len, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
                return e;
            if ((e = hipEventRecord(h->pin_ev[c & 1], h->stream)) != hipSuccess) return e;
        }
        if (c > 0) {  // chunk c-1 out of its buffer while chunk c is in flight
            const size_t p = c - 1, off = p * PIN_CHUNK, len = std::min(PIN_CHUNK, bytes - off);
            if ((e = hipEventSynchronize(h->pin_ev[p & 1])) != hipSuccess) return e;
            par_memcpy(d + off, reinterpret_cast<const char *>(h->pin[p & 1]), len);
        }
    }
    return hipSuccess;
}

// Host-to-device copy of `bytes` from caller (pageable) memory, the mirror
// image: host threads copy chunk c into a pinned buffer (once the DMA that
// last read that buffer is done) while the DMA of chunk c-1 runs.  Queued on
// the handle's stream, so kernels queued after it see the data.
static hipError_t h2d_staged(fh_reconstructor *h, void *dst, const void *src, size_t bytes)
{
    hipError_t e;
    if (bytes < (16ull << 20)) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream);
    if ((e = ensure_pinned(h)) != hipSuccess) return e;
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    const size_t nchunk = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
    for (size_t c = 0; c < nchunk; c++) {
        const size_t off = c * PIN_CHUNK, len = std::min(PIN_CHUNK, bytes - off);
        if ((e = hipEventSynchronize(h->pin_ev[c & 1])) != hipSuccess) return e;  // buffer free
        par_memcpy(reinterpret_cast<char *>(h->pin[c & 1]), s + off, len);
        if ((e = hipMemcpyAsync(d + off, h->pin[c & 1], len, hipMemcpyHostToDevice, h->stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(h->pin_ev[c & 1], h->stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

static void free_chunk(fh_reconstructor *h)
{
    for (float **p : {&h->imgs, &h->means, &h->rots, &h->wts, &h->dfs, &h->dfs2, &h->t3d, &h->wq, &h->rw})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (h->slices) { (void)hipFree(h->slices); h->slices = nullptr; }
    if (h->inter) { (void)hipFree(h->inter); h->inter = nullptr; }
    h->inter_n = 0;
    h->cap_n = h->cap_B = 0;
}

static void free_bins(fh_reconstructor *h)
{
    BinBuffers &b = h->bins;
    for (int **p : {&b.counts, &b.offsets, &b.item_off, &b.tile, &b.list, &b.ctl})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    if (b.items) { (void)hipFree(b.items); b.items = nullptr; }
    if (b.recs) { (void)hipFree(b.recs); b.recs = nullptr; }
    b.rec_cap = 0;
    if (b.part) { (void)hipFree(b.part); b.part = nullptr; }
    b.part_cap = 0;
    h->bins_cap_slices = 0;
}

static void free_ref(fh_reconstructor *h)
{
    if (h->ref) { (void)hipFree(h->ref); h->ref = nullptr; }
}

static void free_volume(fh_reconstructor *h)
{
    if (h->data) { (void)hipFree(h->data); h->data = nullptr; }
    if (h->norm) { (void)hipFree(h->norm); h->norm = nullptr; }
    h->ready = false;
}

// Particles of box P the column-first prep intermediate holds (the prep runs
// in sub-batches of at most this many).
constexpr int PREP_INTER_PARTICLES = 2048;
static size_t prep_inter_bytes(int P, int n)
{
    return sizeof(float2) * (size_t)n * (P / 2 + 1) * prep_inter_pitch(P);
}

static int ensure_chunk(fh_reconstructor *h, int n, int B)
{
    if (n <= h->cap_n && B <= h->cap_B) return FH_OK;
    free_chunk(h);
    const int P = h->P, H = h->H;
    const size_t ns = (size_t)n * h->nsym;
    FH_TRY(h, hipMalloc(&h->imgs, sizeof(float) * (size_t)n * B * B));
    FH_TRY(h, hipMalloc(&h->means, sizeof(float) * n));
    FH_TRY(h, hipMalloc(&h->rots, sizeof(float) * 9 * (size_t)n));
    FH_TRY(h, hipMalloc(&h->wts, sizeof(float) * n));
    FH_TRY(h, hipMalloc(&h->dfs, sizeof(float) * n));
    FH_TRY(h, hipMalloc(&h->dfs2, sizeof(float) * n));
    FH_TRY(h, hipMalloc(&h->slices, sizeof(float2) * (size_t)n * P * (H + 1)));
    if (prep_column_first(P)) {
        h->inter_n = std::min(n, PREP_INTER_PARTICLES);
        FH_TRY(h, hipMalloc(&h->inter, prep_inter_bytes(P, h->inter_n)));
    }
    FH_TRY(h, hipMalloc(&h->t3d, sizeof(float) * 9 * ns));
    FH_TRY(h, hipMalloc(&h->wq, sizeof(float) * ns));
    if (h->usessnr) FH_TRY(h, hipMalloc(&h->rw, sizeof(float) * (size_t)n * (H + 1)));
    h->cap_n = n;
    h->cap_B = B;
    return FH_OK;
}

// Particles per batched launch: the tuned value, or (chunk 0) as many as
// fit in AUTO_CHUNK_BYTES of image staging + slices, within [256, 65536].
// Larger chunks put more slices into each work item, so fewer partial
// bricks are written and merged per particle.
constexpr long long AUTO_CHUNK_BYTES = 48LL << 30;
constexpr size_t WIDE_LDS_PER_CU = 160u << 10;  // LDS of one gfx950 compute unit
static int chunk_for(const fh_tuning &t, int P, int box, long long extra = 0)
{
    if (t.chunk > 0) return t.chunk;
    const long long per = 8LL * P * (P / 2 + 1) + 4LL * box * box + extra;
    return (int)std::max(256LL, std::min(65536LL, AUTO_CHUNK_BYTES / per));
}

// The brick grid of an inserter mode: gauss_2 and trilinear (stencil origin
// floor(p), origins -P/2-1 .. P/2, 8^3 origins per brick), or for the wide
// stencils of S cells per axis 4^3 origins per brick over the origins from
// the lowest one whose stencil reaches a cell in [-P/2, P/2] and that
// wide_setup accepts, up to P/2 (higher origins reach no cell).
static BrickGeom brick_geom(int P, int mode)
{
    BrickGeom g;
    const int H = P / 2, S = wide_cells(mode);
    if (S == 0) {
        g.bs = BRICK;
        g.halo = HALO;
        g.ow = -H - 1;
        g.po = (float)(-H - 1);
    } else {
        int ow;
        float off;  // origin(p) = floor(p - off)
        switch (mode) {
        case 5: ow = -H - 2; off = 0.5f; break;  // gauss_3
        case 6: ow = -H - 4; off = 2.0f; break;  // kaiser_bessel
        case 7: ow = -H - 4; off = 1.5f; break;  // gridding_5
        case 8: ow = -H - 6; off = 2.5f; break;  // gridding_7
        case 9: ow = -H - 7; off = 4.0f; break;  // kaiser_bessel_derived
        default: ow = -H - 4; off = 2.5f; break; // gauss_var, gauss_5, fourier_iter
        }
        g.bs = WBS;
        g.halo = WBS + S - 1;
        g.ow = ow;
        g.po = (float)ow + off;
    }
    g.nb = (H - g.ow + 1 + g.bs - 1) / g.bs;
    g.part_floats = 3 * g.halo * g.halo * g.halo;
    g.per_brick = S != 0;
    g.nbricks = g.nb * g.nb * g.nb;
    g.nkeys = 3 * g.nbricks;
    return g;
}

static int ensure_bins(fh_reconstructor *h, long long nslices)
{
    if (nslices <= h->bins_cap_slices) return FH_OK;
    free_bins(h);
    const BrickGeom g = brick_geom(h->P, h->mode);
    h->geom = g;
    BinBuffers &b = h->bins;
    // at most 5 bricks per brick column per slice (4 shown, see k_bin_cols)
    b.list_cap = 5LL * g.nb * g.nb * nslices;
    b.item_cap = b.list_cap / std::max(1, h->tune.slices_per_item) + g.nkeys + 1;
    FH_TRY(h, hipMalloc(&b.counts, sizeof(int) * g.nkeys));
    FH_TRY(h, hipMalloc(&b.offsets, sizeof(int) * (g.nkeys + 1)));
    FH_TRY(h, hipMalloc(&b.item_off, sizeof(int) * (g.nkeys + 1)));
    FH_TRY(h, hipMalloc(&b.tile, sizeof(int) * 2 * scan_tiles(g)));
    FH_TRY(h, hipMalloc(&b.list, sizeof(int) * b.list_cap));
    FH_TRY(h, hipMalloc(&b.items, sizeof(WorkItem) * b.item_cap));
    FH_TRY(h, hipMalloc(&b.ctl, sizeof(int) * 8));
    FH_TRY(h, hipMalloc(&b.recs, sizeof(SliceRec) * 3 * (size_t)nslices));
    b.rec_cap = nslices;
    h->bins_cap_slices = nslices;
    return FH_OK;
}

// Partial bricks for `items` work items (grown on demand; 8.5 KiB each for
// gauss_2 and gauss_var).
static int ensure_partials(fh_reconstructor *h, long long items)
{
    BinBuffers &b = h->bins;
    if (items <= b.part_cap) return FH_OK;
    if (b.part) { (void)hipFree(b.part); b.part = nullptr; }
    b.part_cap = 0;
    const long long cap = items + items / 4 + 1024;
    FH_TRY(h, hipMalloc(&b.part, sizeof(float) * h->geom.part_floats * (size_t)cap));
    b.part_cap = cap;
    return FH_OK;
}

static InsertConst iconst(const fh_reconstructor *h)
{
    InsertConst c;
    c.P = h->P;
    c.H = h->H;
    c.Pf = (float)h->P;
    c.h = 1.0f / (float)(4.0 / (M_PI * M_PI));  // 1.0f/EMConsts::I2G
    c.mode = h->mode == 10 && h->ref ? 11 : h->mode;
    c.corners = h->corners;
    c.rcut = h->mode == 10 ? h->H + 3 : h->H;
    c.ref = h->mode == 10 ? h->ref : nullptr;
    return c;
}

// ====================================================================== C ABI
// InsertConst::mode of an inserter name (0 for gauss_2 and unknown names)
static int inserter_index(const char *m)
{
    static const char *names[] = {"gauss_2", "trilinear", "gauss_var", "gauss_5", "nearest_neighbor", "gauss_3",
                                  "kaiser_bessel", "gridding_5", "gridding_7", "kaiser_bessel_derived",
                                  "fourier_iter"};
    for (int i = 0; i < 11; i++)
        if (!strcmp(m, names[i])) return i;
    return 0;
}

// Device bytes of a handle of size P and inserter `mode` (a mode name; unknown
// names count as gauss_2) after setup, with batches of `chunk` particles.
long long footprint_bytes_for_mode(int P, int nsym, int chunk, const char *mode_name)
{
    const int mode = mode_name ? inserter_index(mode_name) : 0;
    if (chunk <= 0) chunk = chunk_for(fh_tuning{1, 1, 0, 1, 0}, P, P);
    const long long v = 12LL * P * P * (P / 2 + 1);
    const long long sl = 8LL * chunk * P * (P / 2 + 1) + 4LL * chunk * P * P;
    const BrickGeom g = brick_geom(P, mode);
    const long long nslices = (long long)chunk * nsym;
    const long long list = 5LL * g.nb * g.nb * nslices;
    const long long items = list / 64 + g.nkeys + 1;
    const long long bins = 4LL * list + 16LL * items + 16LL * g.nkeys + 3LL * (long long)sizeof(SliceRec) * nslices;
    // partial bricks: about 1.1 (P/bs)^2 (slice, brick) pairs per slice, one
    // item per 64 pairs of a key plus one per key the slices reach
    const double pb = (double)P / g.bs;
    const long long part_items = (long long)(1.1 * pb * pb * nslices / 64) + g.nkeys / 2;
    const long long inter = prep_column_first(P) ? (long long)prep_inter_bytes(P, std::min(chunk, PREP_INTER_PARTICLES)) : 0;
    return v + sl + inter + bins + 4LL * g.part_floats * part_items;
}

extern "C" {

const char *fh_last_error(const fh_reconstructor *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int fh_create(const fh_params *p, fh_reconstructor **out)
{
    if (!p || !out) {
        g_create_error = "NullPointerException: null argument to fh_create";
        return FH_ERR_NULL;
    }
    *out = nullptr;
    if (p->nx != p->ny || p->ny != p->nz) {
        g_create_error = "ImageDimensionException: fourier_hip requires a cubic size";
        return FH_ERR_IMAGE_DIM;
    }
    if (p->nx < 8 || p->nx % 2 || !fft_size_ok(p->nx)) {
        g_create_error = "ImageDimensionException: size must be even, >= 8 and 13-smooth (2^a 3^b 5^c 7^d 11^e 13^f, got " +
                         std::to_string(p->nx) + ")";
        return FH_ERR_IMAGE_DIM;
    }
    if (p->nx > 1024) {
        g_create_error = "ImageDimensionException: size above 1024 is not supported";
        return FH_ERR_IMAGE_DIM;
    }
    std::string mode = p->mode ? p->mode : "gauss_2";
    int imode;
    if (mode == "gauss_2") imode = 0;
    else if (mode == "trilinear") imode = 1;
    else if (mode == "gauss_var") imode = 2;
    else if (mode == "gauss_5") imode = 3;
    else if (mode == "nearest_neighbor") imode = 4;
    else if (mode == "gauss_3") imode = 5;
    else if (mode == "kaiser_bessel") {
        imode = 6;
        // its window sqrt(1 - (r / (P/2 - 1) - 1)^2) of squared distances r <= 27 is NaN for
        // P/2 - 1 < 13.5: refused here where the reference fills the volume with NaN
        if (p->nx < 32) {
            g_create_error = "InvalidParameterException: kaiser_bessel needs size >= 32";
            return FH_ERR_INVALID_PARAM;
        }
    }
    else if (mode == "fourier_iter") {
        // the fourier_iter reconstructor (libEM/reconstructor.cpp:428-682): its own insertion,
        // no inserter modes, class_ssnr, corners or sqrtnorm
        if (p->usessnr || p->corners || p->sqrtnorm) {
            g_create_error = "InvalidParameterException: fourier_iter takes no usessnr, corners or sqrtnorm";
            return FH_ERR_INVALID_PARAM;
        }
        imode = 10;
    }
    else if (mode == "gridding_5") imode = 7;
    else if (mode == "gridding_7") imode = 8;
    else if (mode == "kaiser_bessel_derived") imode = 9;
    else {
        g_create_error = "NotExistingObjectException: inserter mode '" + mode +
                         "' (fourier_hip supports gauss_2, trilinear, gauss_var, gauss_5, gauss_3, nearest_neighbor, "
                         "kaiser_bessel, kaiser_bessel_derived, gridding_5, gridding_7)";
        return FH_ERR_NOT_EXISTING;
    }
    if (p->usessnr && p->corners) {
        // with corners the reference indexes class_ssnr up to sqrt(2) times past its end
        // (reconstructor.cpp:1171-1175)
        g_create_error = "InvalidParameterException: usessnr with corners is not supported by fourier_hip";
        return FH_ERR_INVALID_PARAM;
    }
    std::string sym = p->sym ? p->sym : "c1";
    std::vector<float> ops;
    if (!sym_ops(sym, ops)) {
        g_create_error = "NotExistingObjectException: symmetry '" + sym + "' (fourier_hip supports c<n>, d<n>, h<n>, tet, oct, icos)";
        return FH_ERR_NOT_EXISTING;
    }
    fh_reconstructor *h = new fh_reconstructor();
    h->P = p->nx;
    h->H = p->nx / 2;
    h->device = p->device;
    h->sym = sym;
    h->mode_name = mode;
    h->mode = imode;
    h->corners = p->corners;
    h->usessnr = p->usessnr ? 1 : 0;
    h->sqrtnorm = p->sqrtnorm;
    h->verbose = p->verbose;
    h->syms = ops;
    h->nsym = (int)ops.size() / 9;
    hipError_t e;
    if ((e = hipSetDevice(h->device)) != hipSuccess || (e = hipStreamCreate(&h->stream)) != hipSuccess) {
        g_create_error = std::string("device error: ") + hipGetErrorString(e);
        delete h;
        return FH_ERR_DEVICE;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && cus > 0) h->cus = cus;
    for (auto &ev : h->ev) (void)hipEventCreate(&ev);
    // constant tables: fast_exp (util.cpp:792-794), fl(i/P), twiddles, syms
    const int P = h->P, H = h->H;
    std::vector<float> et(1000), rt(P + 2);
    for (int i = 0; i < 1000; i++) et[i] = (float)exp(-i / 50.0);
    for (int i = 0; i <= P + 1; i++) rt[i] = (float)(i - H) / (float)P;  // fl(i/P), i in [-H, H+1]
    std::vector<float2> tw(P);
    for (int m = 0; m < P; m++) {
        double a = -2.0 * M_PI * (double)m / P;
        tw[m] = make_float2((float)cos(a), (float)sin(a));
    }
    make_plan(P, &h->plan);
    h->scale = (float)sqrt((double)(1.0f / (float)P * (float)(P + 2)));
    if ((e = hipMalloc(&h->exptab, sizeof(float) * 1000)) != hipSuccess ||
        (e = hipMalloc(&h->rtab, sizeof(float) * (P + 2))) != hipSuccess ||
        (e = hipMalloc(&h->tw, sizeof(float2) * P)) != hipSuccess ||
        (e = hipMalloc(&h->syms_d, sizeof(float) * ops.size())) != hipSuccess ||
        (e = hipMemcpy(h->exptab, et.data(), sizeof(float) * 1000, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->rtab, rt.data(), sizeof(float) * (P + 2), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->tw, tw.data(), sizeof(float2) * P, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->syms_d, ops.data(), sizeof(float) * ops.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        g_create_error = std::string("device error: ") + hipGetErrorString(e);
        fh_destroy(h);
        return FH_ERR_DEVICE;
    }
    *out = h;
    return FH_OK;
}

int fh_destroy(fh_reconstructor *h)
{
    if (!h) return FH_OK;
    {
        std::lock_guard<std::mutex> lk(h->mu);
        (void)hipSetDevice(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        free_chunk(h);
        free_bins(h);
        free_volume(h);
        free_ref(h);
        if (h->pvol) (void)hipFree(h->pvol);
        if (h->ximgs) (void)hipFree(h->ximgs);
        if (h->ximgs2) (void)hipFree(h->ximgs2);
        if (h->xfd) (void)hipFree(h->xfd);
        for (int i = 0; i < 2; i++) {
            if (h->pin[i]) (void)hipHostFree(h->pin[i]);
            if (h->pin_ev[i]) (void)hipEventDestroy(h->pin_ev[i]);
        }
        if (h->partial) (void)hipFree(h->partial);
        if (h->rw1) (void)hipFree(h->rw1);
        if (h->acc8) (void)hipFree(h->acc8);
        for (void *p : {(void *)h->exptab, (void *)h->rtab, (void *)h->tw, (void *)h->syms_d})
            if (p) (void)hipFree(p);
        for (auto &ev : h->ev)
            if (ev) (void)hipEventDestroy(ev);
        if (h->stream) (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return FH_OK;
}

static int setup_locked(fh_reconstructor *h)
{
    FH_TRY(h, hipSetDevice(h->device));
    free_ref(h);  // fourier_iter's setup() drops the reference volume (reconstructor.cpp:462-464)
    const size_t n = vol_elems(h->P);
    if (!h->data) {
        hipError_t e = hipMalloc(&h->data, sizeof(float2) * n);
        if (e != hipSuccess) return fail(h, FH_ERR_NOMEM, std::string("volume allocation: ") + hipGetErrorString(e));
        e = hipMalloc(&h->norm, sizeof(float) * n);
        if (e != hipSuccess) {
            (void)hipFree(h->data);
            h->data = nullptr;
            return fail(h, FH_ERR_NOMEM, std::string("weight allocation: ") + hipGetErrorString(e));
        }
    }
    FH_TRY(h, launch_fill_f32((float *)h->data, 0.f, 2 * n, h->stream));
    FH_TRY(h, launch_fill_f32(h->norm, 0.f, n, h->stream));
    h->ready = true;
    if (h->verbose)
        printf("3D Fourier dimensions are %d %d %d\nYou will require approximately %1.3g GB of memory to reconstruct this volume\n",
               h->P + 2, h->P, h->P, (double)(sizeof(float) * 3 * n) / 1e9);
    return FH_OK;
}

int fh_setup(fh_reconstructor *h)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    return setup_locked(h);
}

// fourier_iter's setup_seed with a complex seed (reconstructor.cpp:482-486):
// the volume cleared and the seed kept as the reference ("we assume tocorner
// was called properly"); the seed weight is not used
static int iter_seed_complex(fh_reconstructor *h, const float *seed)
{
    int rc = setup_locked(h);
    if (rc) return rc;
    const size_t n = vol_elems(h->P);
    FH_TRY(h, hipMalloc(&h->ref, sizeof(float2) * n));
    FH_TRY(h, hipMemcpyAsync(h->ref, seed, sizeof(float2) * n, hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_setup_seed_map(fh_reconstructor *h, const float *map)
{
    if (!h) return FH_ERR_NULL;
    if (!map) return fail(h, FH_ERR_NULL, "NullPointerException: seed map is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->mode != 10)
        return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: a real seed map seeds fourier_iter only");
    int rc = setup_locked(h);
    if (rc) return rc;
    const int P = h->P;
    // the map goes through the particle staging (P images of P x P) and the prep's FFT passes
    if ((rc = ensure_chunk(h, std::max(P, h->cap_n), std::max(P, h->cap_B)))) return rc;
    FH_TRY(h, hipMalloc(&h->ref, sizeof(float2) * vol_elems(P)));
    FH_TRY(h, h2d_staged(h, h->imgs, map, sizeof(float) * (size_t)P * P * P));
    FH_TRY(h, launch_fill_f32(h->means, 0.f, P, h->stream));
    FH_TRY(h, launch_fft3d_forward_corner(h->imgs, P, h->means, h->ref, h->plan, h->tw, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_setup_seed(fh_reconstructor *h, const float *seed, float seed_weight)
{
    if (!h) return FH_ERR_NULL;
    if (!seed) return fail(h, FH_ERR_NULL, "NullPointerException: seed volume is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->mode == 10) return iter_seed_complex(h, seed);
    int rc = setup_locked(h);
    if (rc) return rc;
    const size_t n = vol_elems(h->P);
    FH_TRY(h, hipMemcpyAsync(h->data, seed, sizeof(float2) * n, hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, launch_scale_f32((float *)h->data, seed_weight, 2 * n, h->stream));
    FH_TRY(h, launch_seed_weights(h->norm, h->P, seed_weight, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_setup_seedandweights(fh_reconstructor *h, const float *seed, const float *weights)
{
    if (!h) return FH_ERR_NULL;
    if (!seed || !weights) return fail(h, FH_ERR_NULL, "NullPointerException: seed or weight volume is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->mode == 10) return iter_seed_complex(h, seed);  // fourier_iter ignores the weights (reconstructor.cpp:491-499)
    int rc = setup_locked(h);
    if (rc) return rc;
    const size_t n = vol_elems(h->P);
    FH_TRY(h, hipMemcpyAsync(h->data, seed, sizeof(float2) * n, hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, hipMemcpyAsync(h->norm, weights, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_clear(fh_reconstructor *h)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "clear() before setup()");
    const size_t n = vol_elems(h->P);
    FH_TRY(h, hipSetDevice(h->device));
    free_ref(h);  // fourier_iter's clear() is setup() (reconstructor.cpp:502-504)
    FH_TRY(h, launch_fill_f32((float *)h->data, 0.f, 2 * n, h->stream));
    FH_TRY(h, launch_fill_f32(h->norm, 0.f, n, h->stream));
    return FH_OK;
}

// prep kernels of n particles (box B) into h->slices; dfd: device defocus or null
static hipError_t prep_kernels(fh_reconstructor *h, int n, int B, const float *dev_imgs, const float *dfd,
                               const CtfConst &cc, const float *xfd = nullptr, const float *dfd2 = nullptr)
{
    hipError_t e;
    if ((e = launch_edge_mean(dev_imgs, n, B, h->means, h->stream)) != hipSuccess) return e;
    if (xfd && dfd) {
        // the reference's order (e2make3dpar.py:698-716): the padded image is
        // flipped first -- FFT, CTF_SIGN (times the second sign), inverse FFT
        // -- and preprocess_slice then resamples the flipped image with the
        // 2-D part: the whole prep with the flip, back to real space (phase
        // origin to the centre, 1 / (P^2 scale)), the resampling, the prep
        // again without flip
        const int P = h->P;
        if ((e = launch_prep2d(dev_imgs, n, B, P, h->means, h->slices, h->inter, h->inter_n, dfd, dfd2, cc, h->scale,
                               h->plan, h->tw, h->stream)) != hipSuccess)
            return e;
        const float inv = 1.0f / ((float)P * (float)P * h->scale);
        if ((e = launch_ift2d_batch(h->slices, n, P, h->ximgs, inv, h->plan, h->tw, h->stream)) != hipSuccess) return e;
        if ((e = launch_fill_f32(h->means, 0.f, n, h->stream)) != hipSuccess) return e;
        if ((e = launch_xform2d(h->ximgs, n, P, P, h->means, xfd, h->ximgs2, h->stream)) != hipSuccess) return e;
        return launch_prep2d(h->ximgs2, n, P, P, h->means, h->slices, h->inter, h->inter_n, nullptr, nullptr, cc,
                             h->scale, h->plan, h->tw, h->stream);
    }
    if (xfd) {
        // pad + 2-D resampling into P x P images, then the prep of those (zero mean)
        if ((e = launch_xform2d(dev_imgs, n, B, h->P, h->means, xfd, h->ximgs, h->stream)) != hipSuccess) return e;
        if ((e = launch_fill_f32(h->means, 0.f, n, h->stream)) != hipSuccess) return e;
        return launch_prep2d(h->ximgs, n, h->P, h->P, h->means, h->slices, h->inter, h->inter_n, nullptr, nullptr, cc,
                             h->scale, h->plan, h->tw, h->stream);
    }
    return launch_prep2d(dev_imgs, n, B, h->P, h->means, h->slices, h->inter, h->inter_n, dfd, dfd2, cc, h->scale,
                         h->plan, h->tw, h->stream);
}

// staging of n 2-D transformed particles (P x P) and their inverse transforms;
// two: also the second P x P buffer of the flip-then-resample order
static int ensure_ximgs(fh_reconstructor *h, int n, bool two = false)
{
    if (two && n > h->ximgs2_cap) {
        if (h->ximgs2) (void)hipFree(h->ximgs2);
        h->ximgs2 = nullptr;
        h->ximgs2_cap = 0;
        FH_TRY(h, hipMalloc(&h->ximgs2, sizeof(float) * (size_t)n * h->P * h->P));
        h->ximgs2_cap = n;
    }
    if (n <= h->ximgs_cap) return FH_OK;
    if (h->ximgs) (void)hipFree(h->ximgs);
    if (h->xfd) (void)hipFree(h->xfd);
    h->ximgs = nullptr;
    h->xfd = nullptr;
    h->ximgs_cap = 0;
    FH_TRY(h, hipMalloc(&h->ximgs, sizeof(float) * (size_t)n * h->P * h->P));
    FH_TRY(h, hipMalloc(&h->xfd, sizeof(float) * 6 * (size_t)n));
    h->ximgs_cap = n;
    return FH_OK;
}

// prep of n particles already in h->imgs (box B) into h->slices; xf_host:
// their inverse 2-D transforms (n x 6) or null when all are identities
// defocus2_host: the second defocus of the two-sign flip (or null)
static int prep_chunk(fh_reconstructor *h, int n, int B, const float *dev_imgs, const float *defocus_host,
                      const fh_ctf *ctf, const float *xf_host = nullptr, const float *defocus2_host = nullptr)
{
    const float *dfd = nullptr, *dfd2 = nullptr;
    CtfConst cc{};
    if (defocus_host && ctf) {
        FH_TRY(h, hipMemcpyAsync(h->dfs, defocus_host, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
        dfd = h->dfs;
        cc = ctf_constants(*ctf, h->P);
        if (defocus2_host) {
            FH_TRY(h, hipMemcpyAsync(h->dfs2, defocus2_host, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
            dfd2 = h->dfs2;
        }
    }
    const float *xfd = nullptr;
    if (xf_host) {
        int rc = ensure_ximgs(h, n, dfd != nullptr);
        if (rc) return rc;
        FH_TRY(h, hipMemcpyAsync(h->xfd, xf_host, sizeof(float) * 6 * n, hipMemcpyHostToDevice, h->stream));
        xfd = h->xfd;
    }
    FH_TRY(h, prep_kernels(h, n, B, dev_imgs, dfd, cc, xfd, dfd2));
    if (defocus2_host) FH_TRY(h, hipStreamSynchronize(h->stream));  // the caller's second-defocus array is transient
    return FH_OK;
}

// preprocess_slice of one raw P x P image already in h->imgs (no edge-mean
// step): 2-D transform when xf6 is given, phase origin, r2c FFT, scale
static int prep_single_raw(fh_reconstructor *h, const float *xf6)
{
    const int P = h->P;
    FH_TRY(h, launch_fill_f32(h->means, 0.f, 1, h->stream));
    const float *src = h->imgs;
    if (xf6) {
        int rc = ensure_ximgs(h, 1);
        if (rc) return rc;
        FH_TRY(h, hipMemcpyAsync(h->xfd, xf6, sizeof(float) * 6, hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, launch_xform2d(h->imgs, 1, P, P, h->means, h->xfd, h->ximgs, h->stream));
        src = h->ximgs;
    }
    FH_TRY(h, launch_prep_rows(src, 1, P, P, h->means, h->slices, h->plan, h->tw, h->stream));
    FH_TRY(h, launch_prep_cols(1, P, P, h->slices, nullptr, nullptr, CtfConst{}, h->scale, h->plan, h->tw, h->stream));
    return FH_OK;
}

// Graph form of the batched insertion for device-resident particles: per
// chunk, the host data (rotations, weights, defocus) is copied first, then
// the whole launch sequence -- prep, compose, binning, the persistent
// scatter and the merge -- is captured into one hipGraph and launched.  No
// host read sits inside: the partial-brick buffer holds the binning's
// worst-case item count, and the scatter takes its item count on the
// device.  The captured sequence holds kernels only (no memset nodes).
static int insert_batch_graph(fh_reconstructor *h, const float *imgs, int n, int box, const std::vector<float> &rot9,
                              const float *weights, const float *defocus, const fh_ctf *ctf, int chunk,
                              const float *rw_host, const float *xf_host, const float *df2_host)
{
    int rc = ensure_partials(h, h->bins.item_cap);
    if (rc) return rc;
    if (xf_host && (rc = ensure_ximgs(h, chunk, defocus && ctf))) return rc;
    const InsertConst c = iconst(h);
    fh_stats st{};
    const int wpb = h->tune.waves_per_block;
    const int grid = h->cus * h->tune.blocks_per_cu;
    const CtfConst cc = (defocus && ctf) ? ctf_constants(*ctf, h->P) : CtfConst{};
    FH_TRY(h, hipEventRecord(h->ev[0], h->stream));
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int m = std::min(chunk, n - c0);
        const float *dimgs = imgs + (size_t)c0 * box * box;
        FH_TRY(h, hipMemcpyAsync(h->rots, &rot9[9 * (size_t)c0], sizeof(float) * 9 * m, hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, hipMemcpyAsync(h->wts, weights + c0, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
        if (rw_host)
            FH_TRY(h, hipMemcpyAsync(h->rw, rw_host + (size_t)c0 * (h->H + 1), sizeof(float) * (size_t)m * (h->H + 1),
                                     hipMemcpyHostToDevice, h->stream));
        if (xf_host)
            FH_TRY(h, hipMemcpyAsync(h->xfd, xf_host + 6 * (size_t)c0, sizeof(float) * 6 * m, hipMemcpyHostToDevice,
                                     h->stream));
        const float *dfd = nullptr, *dfd2 = nullptr;
        if (defocus && ctf) {
            FH_TRY(h, hipMemcpyAsync(h->dfs, defocus + c0, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
            dfd = h->dfs;
            if (df2_host) {
                FH_TRY(h, hipMemcpyAsync(h->dfs2, df2_host + c0, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
                dfd2 = h->dfs2;
            }
        }
        FH_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        hipError_t e = prep_kernels(h, m, box, dimgs, dfd, cc, xf_host ? h->xfd : nullptr, dfd2);
        if (e == hipSuccess) e = launch_compose(h->rots, h->wts, m, h->syms_d, h->nsym, h->t3d, h->wq, h->stream);
        if (e == hipSuccess)
            e = launch_bin(h->t3d, h->wq, m * h->nsym, c, h->geom, h->bins, h->tune.slices_per_item, h->stream);
        if (e == hipSuccess)
            e = launch_scatter_insert(h->slices, m * h->nsym, h->nsym, h->t3d, h->wq, c, h->geom, h->bins, h->exptab,
                                      h->rtab, rw_host ? h->rw : nullptr, grid, wpb, h->stream);
        if (e == hipSuccess) e = launch_merge_partials(h->geom, h->bins, c, h->data, h->norm, h->stream);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(h->stream, &g);
        if (e == hipSuccess) e = ee;
        hipGraphExec_t ge = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(ge, h->stream);
        // the counters of this chunk (and the end of its graph) before the next capture
        int ctl[4] = {0, 0, 0, 0};
        if (e == hipSuccess) e = hipMemcpyAsync(ctl, h->bins.ctl, sizeof(ctl), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        FH_TRY(h, e);
        if (ctl[3]) return fail(h, FH_ERR_DEVICE, "internal error: brick list capacity exceeded");
        st.slices += (long long)m * h->nsym;
        st.pairs += ctl[2];
        st.items += ctl[1];
        st.chunks += 1;
    }
    FH_TRY(h, hipEventRecord(h->ev[5], h->stream));
    FH_TRY(h, hipEventSynchronize(h->ev[5]));
    float ms;
    FH_TRY(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[5]));
    st.total_ms = ms;  // per-stage times are not split out in the graph form
    h->stats = st;
    return FH_OK;
}

// rot9: the rotations insert_slice uses (scale, mirror, translation
// stripped).  need_2d: the images still go through preprocess_slice, whose
// 2-D part (translation / scale / mirror) is collected into xf6 (n x 6
// inverse rows; left empty when every particle's is the identity); with
// xf6 null a non-identity 2-D part is refused.
static int check_xforms(fh_reconstructor *h, const float *xforms, int n, int need_2d, std::vector<float> &rot9,
                        std::vector<float> *xf6 = nullptr)
{
    rot9.resize((size_t)9 * n);
    if (xf6) xf6->clear();
    bool any = false;
    std::vector<float> xf;
    if (need_2d && xf6) xf.resize((size_t)6 * n);
    for (int i = 0; i < n; i++) {
        const float *m = xforms + 12 * (size_t)i;
        if (need_2d) {
            if (determinant(m) == 0.0f)
                return fail(h, FH_ERR_INVALID_PARAM, "UnexpectedBehaviorException: particle " + std::to_string(i) +
                                                         ": the determinant of the Transform is 0");
            float inv[6] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
            const bool non_identity = xform2d_inverse(m, inv);
            if (non_identity && !xf6)
                return fail(h, FH_ERR_INVALID_PARAM,
                            "InvalidParameterException: particle " + std::to_string(i) +
                                ": translation/scale/mirror in preprocess_slice is not supported on this path");
            any = any || non_identity;
            if (xf6) memcpy(&xf[6 * (size_t)i], inv, sizeof(inv));
        }
        rotation_only(m, &rot9[9 * (size_t)i]);
    }
    if (xf6 && any) xf6->swap(xf);
    return FH_OK;
}

// The second defocus of the two-sign flip (e2make3dpar.py:698-709,
// e2spa_make3d.py:316-327): EMAN2Ctf's float defocus takes the Python sum
// defocus + dfdf, i.e. the double sum rounded once.  Empty without dfdf.
static void second_defocus(const float *defocus, const fh_ctf *ctf, int n, std::vector<float> &out)
{
    out.clear();
    if (!defocus || !ctf || !ctf->dfdf) return;
    out.resize((size_t)std::max(n, 0));
    for (int i = 0; i < n; i++) out[i] = (float)((double)defocus[i] + (double)ctf->dfdf[i]);
}

int fh_preprocess_slice_xform(fh_reconstructor *h, const float *img, const float xform[12], float *slice_out)
{
    if (!h) return FH_ERR_NULL;
    if (!img || !slice_out) return fail(h, FH_ERR_NULL, "NullPointerException: image is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    std::vector<float> rot9, xf6;
    int rc;
    if (xform && (rc = check_xforms(h, xform, 1, 1, rot9, &xf6))) return rc;
    rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(P, h->cap_B));
    if (rc) return rc;
    FH_TRY(h, hipMemcpyAsync(h->imgs, img, sizeof(float) * P * P, hipMemcpyHostToDevice, h->stream));
    // preprocess_slice has no edge-mean step
    if ((rc = prep_single_raw(h, xf6.empty() ? nullptr : xf6.data()))) return rc;
    FH_TRY(h, hipMemcpyAsync(slice_out, h->slices, sizeof(float2) * P * (H + 1), hipMemcpyDeviceToHost, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_preprocess_slice(fh_reconstructor *h, const float *img, float *slice_out)
{
    return fh_preprocess_slice_xform(h, img, nullptr, slice_out);
}

// plain insertion of one prepared slice (h->slices + idx) with rotation rot9
// one slice of a wide-stencil mode (gauss_var, gauss_5, gauss_3,
// nearest_neighbor, kaiser_bessel(_derived), gridding_5, gridding_7): all its symmetry copies in one LDS-window
// launch (t3d and weights staged in the chunk buffers, which hold >= nsym slices)
static int wide_slice_insert(fh_reconstructor *h, const float2 *slice_d, const float *rot9, float weight,
                             const float *rw_d, const InsertConst &c)
{
    std::vector<float> t(9 * (size_t)h->nsym), wv((size_t)h->nsym, weight);
    for (int s = 0; s < h->nsym; s++) compose9(rot9, &h->syms[9 * s], &t[9 * (size_t)s]);
    FH_TRY(h, hipMemcpyAsync(h->t3d, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, hipMemcpyAsync(h->wq, wv.data(), sizeof(float) * wv.size(), hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, launch_insert_plain_batch(slice_d, h->nsym, h->nsym, h->t3d, h->wq, c, h->data, h->norm, h->exptab,
                                        h->rtab, rw_d, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));  // t and wv are read by the copies above
    return FH_OK;
}

static int insert_slice_plain(fh_reconstructor *h, const float2 *slice_d, const float *rot9, float weight, bool sync_each,
                              const float *rw_d = nullptr)
{
    const InsertConst c = iconst(h);
    if (h->mode >= 2 && !sync_each) return wide_slice_insert(h, slice_d, rot9, weight, rw_d, c);
    for (int s = 0; s < h->nsym; s++) {
        float t[9];
        compose9(rot9, &h->syms[9 * s], t);
        FH_TRY(h, launch_insert_plain(slice_d, c, t, weight, h->data, h->norm, h->exptab, h->rtab, rw_d, h->stream));
        if (sync_each) FH_TRY(h, hipStreamSynchronize(h->stream));
    }
    return FH_OK;
}

// class_ssnr weighting (reconstructor.cpp:1160-1184): the weight of slice
// radius r = hypot_fast_int(x, y) is max(0, ssnr[int(r * sscale)]) with
// sscale = 2 (len - 1) / ny, for r in [0, P/2] (the radius cut keeps r <= P/2).
static void ssnr_radial(int P, const float *ssnr, int len, float *out)
{
    const float iny = (float)P;
    const float sscale = (float)(2.0 * (double)(len - 1) / (double)iny);
    for (int r = 0; r <= P / 2; r++) {
        int i = (int)((float)r * sscale);
        if (i > len - 1) i = len - 1;
        out[r] = std::max(0.0f, ssnr[i]);
    }
}

static int insert_slice_impl(fh_reconstructor *h, const float *data, int is_preprocessed, const float xform[12],
                             float weight, const float *ssnr, int ssnr_len)
{
    if (!h) return FH_ERR_NULL;
    if (!data || !xform) return fail(h, FH_ERR_NULL, "NullPointerException: EMData pointer (input image) is NULL");
    if (ssnr && ssnr_len < 1) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: empty class_ssnr");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "insert_slice() before setup()");
    // usessnr: weight -1 (the SSNR flag) with class_ssnr, else 0 (reconstructor.cpp:1082-1088)
    const bool use_ssnr = h->usessnr && ssnr;
    if (h->usessnr) weight = ssnr ? -1.0f : 0.0f;
    if (weight == 0) return h->mode == 10 ? FH_OK : FH_SKIPPED;  // fourier_iter adds the zeros and returns 0
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    std::vector<float> rot9, xf6;
    int rc = check_xforms(h, xform, 1, !is_preprocessed, rot9, &xf6);
    if (rc) return rc;
    rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(P, h->cap_B));
    if (rc) return rc;
    if (is_preprocessed) {
        FH_TRY(h, hipMemcpyAsync(h->slices, data, sizeof(float2) * P * (H + 1), hipMemcpyHostToDevice, h->stream));
    } else {
        FH_TRY(h, hipMemcpyAsync(h->imgs, data, sizeof(float) * P * P, hipMemcpyHostToDevice, h->stream));
        if ((rc = prep_single_raw(h, xf6.empty() ? nullptr : xf6.data()))) return rc;
    }
    if (use_ssnr) {
        std::vector<float> rw(H + 1);
        ssnr_radial(P, ssnr, ssnr_len, rw.data());
        FH_TRY(h, hipMemcpyAsync(h->rw, rw.data(), sizeof(float) * (H + 1), hipMemcpyHostToDevice, h->stream));
        // the pageable source must stay alive until the copy has been taken
        FH_TRY(h, hipStreamSynchronize(h->stream));
    }
    rc = insert_slice_plain(h, h->slices, rot9.data(), weight, false, use_ssnr ? h->rw : nullptr);
    if (rc) return rc;
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_insert_slice(fh_reconstructor *h, const float *data, int is_preprocessed, const float xform[12], float weight)
{
    return insert_slice_impl(h, data, is_preprocessed, xform, weight, nullptr, 0);
}

int fh_insert_slice_ssnr(fh_reconstructor *h, const float *data, int is_preprocessed, const float xform[12],
                         float weight, const float *class_ssnr, int ssnr_len)
{
    if (h && !class_ssnr) return fail(h, FH_ERR_NULL, "NullPointerException: class_ssnr is NULL");
    return insert_slice_impl(h, data, is_preprocessed, xform, weight, class_ssnr, ssnr_len);
}

static int check_batch_args(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                            const float *weights)
{
    if (!imgs || !xforms || !weights) return fail(h, FH_ERR_NULL, "NullPointerException: batch input is NULL");
    if (n < 0) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: negative particle count");
    if (box < 2 || box > h->P)
        return fail(h, FH_ERR_IMAGE_DIM, "ImageDimensionException: particle box " + std::to_string(box) +
                                             " must be in [2, size=" + std::to_string(h->P) + "]");
    if (!h->ready) return fail(h, FH_ERR_STATE, "insert before setup()");
    return FH_OK;
}

static int insert_batch_impl(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                             const float *weights, const float *defocus, const fh_ctf *ctf, int where,
                             const float *ssnr, int ssnr_len)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->usessnr && !ssnr) {
        // usessnr without class_ssnr: every particle has weight 0 (reconstructor.cpp:1082-1088)
        if (!h->ready) return fail(h, FH_ERR_STATE, "insert before setup()");
        return FH_SKIPPED;
    }
    if (ssnr && ssnr_len < 1) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: empty class_ssnr");
    // with usessnr every particle carries the SSNR flag weight -1 and its radial table
    std::vector<float> rw_host, flag;
    if (h->usessnr) {
        if (!imgs || !xforms) return fail(h, FH_ERR_NULL, "NullPointerException: batch input is NULL");
        rw_host.resize((size_t)std::max(n, 0) * (h->H + 1));
        for (int i = 0; i < n; i++) ssnr_radial(h->P, ssnr + (size_t)i * ssnr_len, ssnr_len, &rw_host[(size_t)i * (h->H + 1)]);
        flag.assign((size_t)std::max(n, 1), -1.0f);
        weights = flag.data();
    }
    const float *rwh = h->usessnr ? rw_host.data() : nullptr;
    int rc = check_batch_args(h, imgs, n, box, xforms, weights);
    if (rc) return rc;
    FH_TRY(h, hipSetDevice(h->device));
    std::vector<float> rot9, xf6;
    rc = check_xforms(h, xforms, n, 1, rot9, &xf6);
    if (rc) return rc;
    const float *xfh = xf6.empty() ? nullptr : xf6.data();
    std::vector<float> df2;
    second_defocus(defocus, ctf, n, df2);
    const float *df2h = df2.empty() ? nullptr : df2.data();
    // a CTF flip before a 2-D resampling stages two P x P images per particle
    const long long xbytes = xfh ? ((defocus && ctf) ? 8LL : 4LL) * h->P * h->P : 0;
    // the brick list (at most 5 nb^2 entries per slice) is indexed with int
    const BrickGeom bg = brick_geom(h->P, h->mode);
    const long long max_slices = std::max(1LL, (long long)INT32_MAX / (5LL * bg.nb * bg.nb) - 1);
    const int chunk = (int)std::max(1LL, std::min({(long long)chunk_for(h->tune, h->P, box, xbytes),
                                                   (long long)std::max(n, 1), max_slices / h->nsym}));
    rc = ensure_chunk(h, chunk, box);
    if (rc) return rc;
    rc = ensure_bins(h, (long long)chunk * h->nsym);
    if (rc) return rc;
    if (h->tune.graph && where == FH_DEVICE && h->mode < 2)
        return insert_batch_graph(h, imgs, n, box, rot9, weights, defocus, ctf, chunk, rwh, xfh, df2h);
    const InsertConst c = iconst(h);
    fh_stats st{};
    float ms;
    const int wpb = h->tune.waves_per_block;
    const int grid = h->cus * h->tune.blocks_per_cu;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int m = std::min(chunk, n - c0);
        const float *src = imgs + (size_t)c0 * box * box;
        const float *dimgs;
        FH_TRY(h, hipEventRecord(h->ev[0], h->stream));
        if (where == FH_DEVICE) {
            dimgs = src;
        } else {
            FH_TRY(h, h2d_staged(h, h->imgs, src, sizeof(float) * (size_t)m * box * box));
            dimgs = h->imgs;
        }
        rc = prep_chunk(h, m, box, dimgs, defocus ? defocus + c0 : nullptr, ctf, xfh ? xfh + 6 * (size_t)c0 : nullptr,
                        df2h ? df2h + c0 : nullptr);
        if (rc) return rc;
        FH_TRY(h, hipMemcpyAsync(h->rots, &rot9[9 * (size_t)c0], sizeof(float) * 9 * m, hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, hipMemcpyAsync(h->wts, weights + c0, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
        if (rwh)
            FH_TRY(h, hipMemcpyAsync(h->rw, rwh + (size_t)c0 * (h->H + 1), sizeof(float) * (size_t)m * (h->H + 1),
                                     hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, launch_compose(h->rots, h->wts, m, h->syms_d, h->nsym, h->t3d, h->wq, h->stream));
        FH_TRY(h, hipEventRecord(h->ev[1], h->stream));
        if (h->mode >= 2 && !wide_cells(h->mode)) {
            // nearest_neighbor (one cell per pixel): the LDS-window tile form
            FH_TRY(h, launch_insert_plain_batch(h->slices, m * h->nsym, h->nsym, h->t3d, h->wq, c, h->data, h->norm,
                                                h->exptab, h->rtab, rwh ? h->rw : nullptr, h->stream));
            FH_TRY(h, hipEventRecord(h->ev[5], h->stream));
            FH_TRY(h, hipEventSynchronize(h->ev[5]));
            FH_TRY(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
            st.prep_ms += ms;
            FH_TRY(h, hipEventElapsedTime(&ms, h->ev[1], h->ev[5]));
            st.insert_kernel_ms += ms;
            st.insert_ms += ms;
            FH_TRY(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[5]));
            st.total_ms += ms;
            st.slices += (long long)m * h->nsym;
            st.chunks += 1;
            continue;
        }
        FH_TRY(h, launch_bin(h->t3d, h->wq, m * h->nsym, c, h->geom, h->bins, h->tune.slices_per_item, h->stream));
        FH_TRY(h, hipEventRecord(h->ev[2], h->stream));
        // the work item count sizes the partial-brick buffer
        int ctl[4];
        FH_TRY(h, hipMemcpyAsync(ctl, h->bins.ctl, sizeof(ctl), hipMemcpyDeviceToHost, h->stream));
        FH_TRY(h, hipStreamSynchronize(h->stream));
        if (ctl[3]) return fail(h, FH_ERR_DEVICE, "internal error: brick list capacity exceeded");
        rc = ensure_partials(h, ctl[1]);
        if (rc) return rc;
        FH_TRY(h, hipEventRecord(h->ev[3], h->stream));
        if (ctl[1] > 0 && wide_cells(h->mode)) {
            // the wide stencils: one wave per item, 4 per workgroup (2 for 9-cell stencils), as
            // many workgroups as fit the LDS of every compute unit
            const int ww = wide_cells(h->mode) > 7 ? 2 : 4;
            const int per_cu = std::max(1, (int)(WIDE_LDS_PER_CU / scatter_wide_lds_bytes(h->mode, ww)));
            FH_TRY(h, launch_scatter_wide(h->slices, m * h->nsym, h->nsym, h->t3d, h->wq, c, h->geom, h->bins,
                                          h->exptab, h->rtab, rwh ? h->rw : nullptr,
                                          std::min(h->cus * per_cu, (ctl[1] + ww - 1) / ww), ww, h->stream));
        } else if (ctl[1] > 0) {
            FH_TRY(h, launch_scatter_insert(h->slices, m * h->nsym, h->nsym, h->t3d, h->wq, c, h->geom, h->bins,
                                            h->exptab, h->rtab, rwh ? h->rw : nullptr,
                                            std::min(grid, (ctl[1] + wpb - 1) / wpb), wpb, h->stream));
        }
        FH_TRY(h, hipEventRecord(h->ev[4], h->stream));
        if (ctl[1] > 0) FH_TRY(h, launch_merge_partials(h->geom, h->bins, c, h->data, h->norm, h->stream));
        FH_TRY(h, hipEventRecord(h->ev[5], h->stream));
        FH_TRY(h, hipEventSynchronize(h->ev[5]));
        FH_TRY(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        st.prep_ms += ms;
        FH_TRY(h, hipEventElapsedTime(&ms, h->ev[1], h->ev[2]));
        st.bin_ms += ms;
        FH_TRY(h, hipEventElapsedTime(&ms, h->ev[3], h->ev[4]));
        st.insert_kernel_ms += ms;
        st.insert_ms += ms;
        FH_TRY(h, hipEventElapsedTime(&ms, h->ev[4], h->ev[5]));
        st.merge_ms += ms;
        st.insert_ms += ms;
        FH_TRY(h, hipEventElapsedTime(&ms, h->ev[0], h->ev[5]));
        st.total_ms += ms;
        st.slices += (long long)m * h->nsym;
        st.pairs += ctl[2];
        st.items += ctl[1];
        st.chunks += 1;
    }
    h->stats = st;
    return FH_OK;
}

int fh_insert_batch(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms, const float *weights,
                    const float *defocus, const fh_ctf *ctf, int where)
{
    return insert_batch_impl(h, imgs, n, box, xforms, weights, defocus, ctf, where, nullptr, 0);
}

int fh_insert_batch_ssnr(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                         const float *weights, const float *class_ssnr, int ssnr_len, const float *defocus,
                         const fh_ctf *ctf, int where)
{
    if (h && !class_ssnr) return fail(h, FH_ERR_NULL, "NullPointerException: class_ssnr is NULL");
    if (h && !h->usessnr)  // usessnr off: class_ssnr is ignored, the weights apply
        return insert_batch_impl(h, imgs, n, box, xforms, weights, defocus, ctf, where, nullptr, 0);
    return insert_batch_impl(h, imgs, n, box, xforms, weights, defocus, ctf, where, class_ssnr, ssnr_len);
}

int fh_prep_batch(fh_reconstructor *h, const float *imgs, int n, int box, const float *defocus, const fh_ctf *ctf,
                  float *slices_out, int where)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (!imgs || !slices_out) return fail(h, FH_ERR_NULL, "NullPointerException: batch input is NULL");
    if (box < 2 || box > h->P)
        return fail(h, FH_ERR_IMAGE_DIM, "ImageDimensionException: particle box must be in [2, size]");
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    const int chunk = std::max(1, std::min(chunk_for(h->tune, P, box), std::max(n, 1)));
    int rc = ensure_chunk(h, chunk, box);
    if (rc) return rc;
    std::vector<float> df2;
    second_defocus(defocus, ctf, n, df2);
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int m = std::min(chunk, n - c0);
        const float *src = imgs + (size_t)c0 * box * box;
        const float *dimgs = src;
        if (where != FH_DEVICE) {
            FH_TRY(h, h2d_staged(h, h->imgs, src, sizeof(float) * (size_t)m * box * box));
            dimgs = h->imgs;
        }
        rc = prep_chunk(h, m, box, dimgs, defocus ? defocus + c0 : nullptr, ctf, nullptr,
                        df2.empty() ? nullptr : df2.data() + c0);
        if (rc) return rc;
        FH_TRY(h, hipMemcpyAsync(slices_out + (size_t)c0 * P * (P + 2), h->slices, sizeof(float2) * (size_t)m * P * (H + 1),
                                 where == FH_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream));
        FH_TRY(h, hipStreamSynchronize(h->stream));
    }
    return FH_OK;
}

static int insert_plain_impl(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                             const float *weights, const float *defocus, const fh_ctf *ctf, int where);

int fh_insert_plain(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms, const float *weights,
                    const float *defocus, const fh_ctf *ctf)
{
    return insert_plain_impl(h, imgs, n, box, xforms, weights, defocus, ctf, FH_HOST);
}

int fh_insert_plain_where(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                          const float *weights, const float *defocus, const fh_ctf *ctf, int where)
{
    return insert_plain_impl(h, imgs, n, box, xforms, weights, defocus, ctf, where);
}

static int insert_plain_impl(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms,
                             const float *weights, const float *defocus, const fh_ctf *ctf, int where)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    int rc = check_batch_args(h, imgs, n, box, xforms, weights);
    if (rc) return rc;
    if (h->usessnr) return FH_SKIPPED;  // no class_ssnr: weight 0 (reconstructor.cpp:1082-1088)
    FH_TRY(h, hipSetDevice(h->device));
    std::vector<float> rot9, xf6;
    rc = check_xforms(h, xforms, n, 1, rot9, &xf6);
    if (rc) return rc;
    std::vector<float> df2;
    second_defocus(defocus, ctf, n, df2);
    rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(box, h->cap_B));
    if (rc) return rc;
    for (int i = 0; i < n; i++) {
        if (weights[i] == 0) continue;
        // per-slice host->device copy (reconstructor.cpp:1074-1078), unless the particle is on the device
        const float *dimg = imgs + (size_t)i * box * box;
        if (where != FH_DEVICE) {
            FH_TRY(h, hipMemcpy(h->imgs, dimg, sizeof(float) * box * box, hipMemcpyHostToDevice));
            dimg = h->imgs;
        }
        rc = prep_chunk(h, 1, box, dimg, defocus ? defocus + i : nullptr, ctf,
                        xf6.empty() ? nullptr : &xf6[6 * (size_t)i], df2.empty() ? nullptr : &df2[i]);
        if (rc) return rc;
        // one launch per symmetry op, synchronised (cuda_reconstructor.cu:171-193)
        rc = insert_slice_plain(h, h->slices, &rot9[9 * (size_t)i], weights[i], true);
        if (rc) return rc;
    }
    return FH_OK;
}

int fh_projection(fh_reconstructor *h, const float xform[12], int ret_fourier, float *out)
{
    if (!h) return FH_ERR_NULL;
    if (!xform || !out) return fail(h, FH_ERR_NULL, "NullPointerException: projection output is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "projection() before setup() or after finish()");
    if (h->mode == 10)
        return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: fourier_iter has no projection()");
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    // only the rotation enters the slice (reconstructor.cpp:1398-1401); the
    // translation (Transform::get_trans, transform.cpp:1077-1092) shifts it
    const bool mirror = determinant(xform) < 0;
    float tx = mirror ? -xform[3] : xform[3], ty = xform[7];
    apply_precision(tx, ERR_LIMIT);
    apply_precision(ty, ERR_LIMIT);
    float rot9[9];
    rotation_only(xform, rot9);
    int rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(P, h->cap_B));
    if (rc) return rc;
    const InsertConst c = iconst(h);
    FH_TRY(h, launch_project(h->data, h->norm, c, rot9, h->rtab, tx, ty, ret_fourier ? 0 : 1, h->slices, h->stream));
    if (ret_fourier) {
        FH_TRY(h, hipMemcpyAsync(out, h->slices, sizeof(float2) * P * (H + 1), hipMemcpyDeviceToHost, h->stream));
    } else {
        FH_TRY(h, launch_ift2d(h->slices, P, h->imgs, h->plan, h->tw, h->stream));
        FH_TRY(h, hipMemcpyAsync(out, h->imgs, sizeof(float) * P * P, hipMemcpyDeviceToHost, h->stream));
    }
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

// rmax of do_compare_slice_work (libEM/reconstructor.cpp:1272-1274)
static float agree_rmax(int P, float apix)
{
    const float iny = (float)P;
    float rmax = (float)(iny * apix / 12.0);
    if (rmax > iny / 2.0 - 1.0) rmax = (float)(iny / 2.0 - 1.0);
    if (rmax < 5.0) rmax = 5.0f;
    return rmax;
}

// The five attributes from the eight sums {dot, vweight, power, power2,
// dotlow, dlpow, dlpow2, nval} (libEM/reconstructor.cpp:1365-1378).
static void agree_outputs(const double *a, float weight, float out[5])
{
    double dot = a[0], vweight = a[1], power = a[2], power2 = a[3], dotlow = a[4], dlpow = a[5], dlpow2 = a[6];
    const int nval = (int)a[7];
    dot /= sqrt(power * power2);
    if (dlpow * dlpow2 > 0) dotlow /= sqrt(dlpow * dlpow2);
    out[0] = (float)(power2 <= 0 ? 1.0 : sqrt(power / power2));
    out[1] = (float)dot;
    out[2] = (float)dotlow;
    const float rw = weight <= 0 ? 1.0f : 1.0f / weight;
    out[3] = (float)(dot * rw / ((rw - 1.0) * dot + 1.0));
    out[4] = (float)(vweight / (float)(nval));
}

// One slice already in slice_d: optional subtraction, compare, re-add
// (libEM/reconstructor.cpp:1212-1257) -> out[5] (reconstructor.cpp:1365-1378).
// rw_d (device, P/2+1 radial weights): the reference's own first step for a
// slice carrying class_ssnr -- do_insert_slice_work with -weight takes its
// weight < 0 SSNR branch and adds the slice with those weights
// (reconstructor.cpp:1160-1184,1236) -- in place of the removal
static int agree_one(fh_reconstructor *h, const float2 *slice_d, const float *rot9, float weight, int sub, float apix,
                     float out[5], const float *rw_d = nullptr)
{
    InsertConst c = iconst(h);
    c.corners = 0;  // do_insert_slice_work default (reconstructor.h:540)
    const int nb = agree_blocks(h->P);
    std::vector<float> t(9 * (size_t)h->nsym);
    for (int s = 0; s < h->nsym; s++) compose9(rot9, &h->syms[9 * s], &t[9 * (size_t)s]);
    FH_TRY(h, hipMemcpyAsync(h->t3d, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, h->stream));
    if (sub && h->mode >= 2) {
        int rc = wide_slice_insert(h, slice_d, rot9, -weight, rw_d, c);
        if (rc) return rc;
    } else if (sub)
        for (int s = 0; s < h->nsym; s++)
            FH_TRY(h, launch_insert_plain(slice_d, c, &t[9 * (size_t)s], -weight, h->data, h->norm, h->exptab, h->rtab, rw_d,
                                          h->stream));
    FH_TRY(h, launch_agree(slice_d, h->nsym, h->t3d, h->nsym, c, agree_rmax(h->P, apix), h->data, h->norm, h->rtab,
                           h->partial, h->stream));
    std::vector<double> part((size_t)h->nsym * nb * 8);
    FH_TRY(h, hipMemcpyAsync(part.data(), h->partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, h->stream));
    if (sub && h->mode >= 2) {
        int rc = wide_slice_insert(h, slice_d, rot9, weight, nullptr, c);
        if (rc) return rc;
    } else if (sub)
        for (int s = 0; s < h->nsym; s++)
            FH_TRY(h, launch_insert_plain(slice_d, c, &t[9 * (size_t)s], weight, h->data, h->norm, h->exptab, h->rtab, nullptr,
                                          h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < part.size(); i++) a[i % 8] += part[i];
    agree_outputs(a, weight, out);
    return FH_OK;
}

static int ensure_partial(fh_reconstructor *h)
{
    if (h->partial) return FH_OK;
    FH_TRY(h, hipMalloc(&h->partial, sizeof(double) * 8 * (size_t)agree_blocks(h->P) * h->nsym));
    return FH_OK;
}

// ---- allocation-free agreement ----
// Scratch of fh_determine_slice_agreement_noalloc: the per-operator
// rotations and weights, the compare kernel's per-block partial sums and
// their eight totals, in one file-scope device array.  Each device holds its
// own copy at its own address (hipGetSymbolAddress, resolved once per device
// with that device current); host code uses only that address, never the
// array's name.  Every call on a device shares the device's copy: calls that
// overlap on one device race.  Nothing in it needs clearing: every slot a
// call reads was written earlier in the same call.
constexpr size_t NOALLOC_HEAD = 4096;             // t3d (64 x 9 floats), weights (64 floats), 8 doubles
constexpr size_t NOALLOC_PART = 2u << 20;         // partial sums: 8 doubles per (block, operator)
constexpr int NOALLOC_MAX_SYM = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_HEAD + NOALLOC_PART];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

// the address of `scratch` on h's device (current), resolved on first use
static int scratch_base(fh_reconstructor *h, unsigned char **out)
{
    if (h->device < 0 || h->device >= 64) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: device index");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        void *ptr = nullptr;
        FH_TRY(h, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(h, FH_ERR_DEVICE, "device error: scratch symbol has no address");
        g_scratch_ptr[h->device] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[h->device]);
    return FH_OK;
}

// the eight sums over n partial records of 8 doubles, in one workgroup
__global__ void __launch_bounds__(256) k_sum_partials8(const double *__restrict__ part, int n, double *__restrict__ out8)
{
    __shared__ double red[256];
    const int k = threadIdx.x & 7, lane = threadIdx.x >> 3;  // 32 strided partial sums per component
    double a = 0.0;
    for (int i = lane; i < n; i += 32) a += part[(size_t)i * 8 + k];
    red[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x < 8) {
        double t = 0.0;
        for (int j = 0; j < 32; j++) t += red[j * 8 + threadIdx.x];
        out8[threadIdx.x] = t;
    }
}

int fh_determine_slice_agreement_noalloc(fh_reconstructor *h, const float *slice_dev, const float xform[12],
                                         float weight, int sub, float apix, float out[5])
{
    if (!h) return FH_ERR_NULL;
    if (!slice_dev || !xform || !out) return fail(h, FH_ERR_NULL, "NullPointerException: slice, xform or out is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "determine_slice_agreement() before setup()");
    if (h->mode == 10) return FH_OK;  // fourier_iter: returns without comparing (reconstructor.cpp:639)
    FH_TRY(h, hipSetDevice(h->device));
    unsigned char *base = nullptr;
    int rc = scratch_base(h, &base);
    if (rc) return rc;
    const int nb = agree_blocks(h->P), ns = h->nsym;
    const size_t need = sizeof(double) * 8 * (size_t)nb * ns;
    if (ns > NOALLOC_MAX_SYM || need > NOALLOC_PART)
        return fail(h, FH_ERR_INVALID_PARAM,
                    "InvalidParameterException: the allocation-free agreement holds at most 2 MiB of partial sums "
                    "(8 doubles per compare block and symmetry operator) and 64 operators; this size and "
                    "symmetry need " + std::to_string(need) + " bytes and " + std::to_string(ns) + " operators");
    float *t3d_d = reinterpret_cast<float *>(base);
    float *wq_d = reinterpret_cast<float *>(base + sizeof(float) * 9 * NOALLOC_MAX_SYM);
    double *sum8_d = reinterpret_cast<double *>(base + sizeof(float) * 10 * NOALLOC_MAX_SYM);
    double *part_d = reinterpret_cast<double *>(base + NOALLOC_HEAD);
    const float2 *slice_d = reinterpret_cast<const float2 *>(slice_dev);
    // host arrays live until the final synchronisation: the copies below read them
    float rot9[9], t[9 * NOALLOC_MAX_SYM], wneg[NOALLOC_MAX_SYM], wpos[NOALLOC_MAX_SYM];
    rotation_only(xform, rot9);
    for (int s = 0; s < ns; s++) {
        compose9(rot9, &h->syms[9 * s], &t[9 * s]);
        wneg[s] = -weight;
        wpos[s] = weight;
    }
    InsertConst c = iconst(h);
    c.corners = 0;  // do_insert_slice_work default (reconstructor.h:540)
    FH_TRY(h, hipMemcpyAsync(t3d_d, t, sizeof(float) * 9 * ns, hipMemcpyHostToDevice, h->stream));
    FH_TRY(h, hipMemcpyAsync(wq_d, wneg, sizeof(float) * ns, hipMemcpyHostToDevice, h->stream));
    if (sub) {
        if (h->mode >= 2)
            FH_TRY(h, launch_insert_plain_batch(slice_d, ns, ns, t3d_d, wq_d, c, h->data, h->norm, h->exptab, h->rtab,
                                                nullptr, h->stream));
        else
            for (int s = 0; s < ns; s++)
                FH_TRY(h, launch_insert_plain(slice_d, c, &t[9 * s], -weight, h->data, h->norm, h->exptab, h->rtab,
                                              nullptr, h->stream));
    }
    FH_TRY(h, launch_agree(slice_d, ns, t3d_d, ns, c, agree_rmax(h->P, apix), h->data, h->norm, h->rtab, part_d,
                           h->stream));
    hipLaunchKernelGGL(k_sum_partials8, dim3(1), dim3(256), 0, h->stream, part_d, nb * ns, sum8_d);
    FH_TRY(h, hipGetLastError());
    double a[8];
    FH_TRY(h, hipMemcpyAsync(a, sum8_d, sizeof(a), hipMemcpyDeviceToHost, h->stream));
    if (sub) {
        // the weights of the re-add, stream-ordered behind the removal's reads of wq_d
        FH_TRY(h, hipMemcpyAsync(wq_d, wpos, sizeof(float) * ns, hipMemcpyHostToDevice, h->stream));
        if (h->mode >= 2)
            FH_TRY(h, launch_insert_plain_batch(slice_d, ns, ns, t3d_d, wq_d, c, h->data, h->norm, h->exptab, h->rtab,
                                                nullptr, h->stream));
        else
            for (int s = 0; s < ns; s++)
                FH_TRY(h, launch_insert_plain(slice_d, c, &t[9 * s], weight, h->data, h->norm, h->exptab, h->rtab,
                                              nullptr, h->stream));
    }
    FH_TRY(h, hipStreamSynchronize(h->stream));
    agree_outputs(a, weight, out);
    return FH_OK;
}

int fh_determine_slice_agreement(fh_reconstructor *h, const float *data, int is_preprocessed, const float xform[12],
                                 float weight, int sub, float apix, float out[5])
{
    if (!h) return FH_ERR_NULL;
    if (!data || !xform || !out) return fail(h, FH_ERR_NULL, "NullPointerException: EMData pointer (input image) is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "determine_slice_agreement() before setup()");
    if (h->mode == 10) return FH_OK;  // fourier_iter: returns without comparing (reconstructor.cpp:639)
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    std::vector<float> rot9, xf6;
    int rc = check_xforms(h, xform, 1, !is_preprocessed, rot9, &xf6);
    if (rc) return rc;
    if ((rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(P, h->cap_B)))) return rc;
    if ((rc = ensure_partial(h))) return rc;
    if (is_preprocessed) {
        FH_TRY(h, hipMemcpyAsync(h->slices, data, sizeof(float2) * P * (H + 1), hipMemcpyHostToDevice, h->stream));
    } else {
        FH_TRY(h, hipMemcpyAsync(h->imgs, data, sizeof(float) * P * P, hipMemcpyHostToDevice, h->stream));
        if ((rc = prep_single_raw(h, xf6.empty() ? nullptr : xf6.data()))) return rc;
    }
    return agree_one(h, h->slices, rot9.data(), weight, sub, apix, out);
}

int fh_determine_slice_agreement_ssnr(fh_reconstructor *h, const float *data, int is_preprocessed, const float xform[12],
                                      float weight, int sub, float apix, const float *class_ssnr, int ssnr_len,
                                      float out[5])
{
    if (!h) return FH_ERR_NULL;
    if (!data || !xform || !out) return fail(h, FH_ERR_NULL, "NullPointerException: EMData pointer (input image) is NULL");
    if (sub && !class_ssnr)
        return fail(h, FH_ERR_NOT_EXISTING, "NotExistingObjectException: The requested key does not exist: class_ssnr");
    if (sub && ssnr_len < 1) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: empty class_ssnr");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "determine_slice_agreement() before setup()");
    if (h->mode == 10) return FH_OK;  // fourier_iter: returns without comparing (reconstructor.cpp:639)
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    std::vector<float> rot9, xf6;
    int rc = check_xforms(h, xform, 1, !is_preprocessed, rot9, &xf6);
    if (rc) return rc;
    if ((rc = ensure_chunk(h, std::max(1, h->cap_n), std::max(P, h->cap_B)))) return rc;
    if ((rc = ensure_partial(h))) return rc;
    if (is_preprocessed) {
        FH_TRY(h, hipMemcpyAsync(h->slices, data, sizeof(float2) * P * (H + 1), hipMemcpyHostToDevice, h->stream));
    } else {
        FH_TRY(h, hipMemcpyAsync(h->imgs, data, sizeof(float) * P * P, hipMemcpyHostToDevice, h->stream));
        if ((rc = prep_single_raw(h, xf6.empty() ? nullptr : xf6.data()))) return rc;
    }
    std::vector<float> rw(H + 1);
    if (sub) {
        if (!h->rw1) FH_TRY(h, hipMalloc(&h->rw1, sizeof(float) * (H + 1)));
        ssnr_radial(P, class_ssnr, ssnr_len, rw.data());
        FH_TRY(h, hipMemcpyAsync(h->rw1, rw.data(), sizeof(float) * (H + 1), hipMemcpyHostToDevice, h->stream));
    }
    // agree_one synchronises the stream before it returns: rw outlives the copy
    return agree_one(h, h->slices, rot9.data(), weight, sub, apix, out, sub ? h->rw1 : nullptr);
}

int fh_agreement_batch(fh_reconstructor *h, const float *imgs, int n, int box, const float *xforms, const float *weights,
                       const float *defocus, const fh_ctf *ctf, int sub, float apix, float *out, int where)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    int rc = check_batch_args(h, imgs, n, box, xforms, weights);
    if (rc) return rc;
    if (!out) return fail(h, FH_ERR_NULL, "NullPointerException: output is NULL");
    if (h->mode == 10) return FH_OK;  // fourier_iter: no comparison (reconstructor.cpp:639); out untouched
    FH_TRY(h, hipSetDevice(h->device));
    // the own-contribution gather knows 2x2x2 stencils only: the other modes
    // take the reference's sequence per particle (remove, compare, re-add)
    const bool seq = h->mode >= 2;
    if (seq && (rc = ensure_partial(h))) return rc;
    std::vector<float> rot9, xf6;
    if ((rc = check_xforms(h, xforms, n, 1, rot9, &xf6))) return rc;
    const float *xfh = xf6.empty() ? nullptr : xf6.data();
    std::vector<float> df2;
    second_defocus(defocus, ctf, n, df2);
    const long long xbytes = xfh ? ((defocus && ctf) ? 8LL : 4LL) * h->P * h->P : 0;
    const int chunk = std::max(1, std::min(chunk_for(h->tune, h->P, box, xbytes), std::max(n, 1)));
    if ((rc = ensure_chunk(h, chunk, box))) return rc;
    if (!h->acc8 || h->acc8_cap < chunk) {
        if (h->acc8) (void)hipFree(h->acc8);
        h->acc8 = nullptr;
        FH_TRY(h, hipMalloc(&h->acc8, sizeof(double) * 8 * (size_t)chunk));
        h->acc8_cap = chunk;
    }
    const InsertConst c = iconst(h);
    const float rmax = agree_rmax(h->P, apix);
    std::vector<float> res(5 * (size_t)std::max(n, 1));
    std::vector<double> acc(8 * (size_t)chunk);
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int m = std::min(chunk, n - c0);
        const float *src = imgs + (size_t)c0 * box * box;
        const float *dimgs = src;
        if (where != FH_DEVICE) {
            FH_TRY(h, h2d_staged(h, h->imgs, src, sizeof(float) * (size_t)m * box * box));
            dimgs = h->imgs;
        }
        if ((rc = prep_chunk(h, m, box, dimgs, defocus ? defocus + c0 : nullptr, ctf, xfh ? xfh + 6 * (size_t)c0 : nullptr,
                             df2.empty() ? nullptr : df2.data() + c0)))
            return rc;
        if (seq) {
            for (int i = 0; i < m; i++)
                if ((rc = agree_one(h, h->slices + (size_t)i * h->P * (h->H + 1), &rot9[9 * (size_t)(c0 + i)],
                                    weights[c0 + i], sub, apix, &res[5 * (size_t)(c0 + i)])))
                    return rc;
            continue;
        }
        FH_TRY(h, hipMemcpyAsync(h->rots, &rot9[9 * (size_t)c0], sizeof(float) * 9 * m, hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, hipMemcpyAsync(h->wts, weights + c0, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
        FH_TRY(h, launch_compose(h->rots, h->wts, m, h->syms_d, h->nsym, h->t3d, h->wq, h->stream));
        FH_TRY(h, launch_fill_i32((int *)h->acc8, 0, 16 * (size_t)m, h->stream));
        // every particle is compared against V minus its own contribution
        // (what e2make3d.py:447-491 gets from determine_slice_agreement called
        // particle by particle), recomputed by a gather instead of mutating V
        FH_TRY(h, launch_agree_batch(h->slices, h->nsym, h->t3d, h->wq, m * h->nsym, c, rmax, sub, h->data, h->norm,
                                     h->exptab, h->rtab, h->acc8, h->stream));
        FH_TRY(h, hipMemcpyAsync(acc.data(), h->acc8, sizeof(double) * 8 * m, hipMemcpyDeviceToHost, h->stream));
        FH_TRY(h, hipStreamSynchronize(h->stream));
        for (int i = 0; i < m; i++)
            agree_outputs(&acc[8 * (size_t)i], weights[c0 + i], &res[5 * (size_t)(c0 + i)]);
    }
    if (n > 0) {
        if (where == FH_DEVICE)
            FH_TRY(h, hipMemcpy(out, res.data(), sizeof(float) * 5 * n, hipMemcpyHostToDevice));
        else
            memcpy(out, res.data(), sizeof(float) * 5 * n);
    }
    return FH_OK;
}

int fh_finish(fh_reconstructor *h, int doift, float *out, float *norm_out)
{
    if (!h) return FH_ERR_NULL;
    if (!out) return fail(h, FH_ERR_NULL, "NullPointerException: output buffer is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "finish() before setup()");
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P;
    const size_t n = vol_elems(P);
    float *planes = nullptr;
    FH_TRY(h, hipMalloc(&planes, sizeof(float) * 2 * P * P));
    hipError_t e = launch_normalize(h->data, h->norm, P, h->sqrtnorm, planes, h->stream);
    if (e == hipSuccess && doift) {
        e = launch_fft3d_inverse_zy(h->data, P, h->plan, h->tw, h->stream);
        float *real = nullptr;
        if (e == hipSuccess) e = hipMalloc(&real, sizeof(float) * (size_t)P * P * P);
        if (e == hipSuccess) {
            const float sc = 1.0f / (float)((size_t)P * P * P);
            e = launch_c2r_x(h->data, P, real, sc, h->plan, h->tw, h->stream);
            if (e == hipSuccess) e = d2h_staged(h, out, real, sizeof(float) * (size_t)P * P * P);
            (void)hipFree(real);
        }
    } else if (e == hipSuccess) {
        e = d2h_staged(h, out, h->data, sizeof(float2) * n);
    }
    if (e == hipSuccess && norm_out) e = d2h_staged(h, norm_out, h->norm, sizeof(float) * n);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(planes);
    if (e != hipSuccess) return fail(h, FH_ERR_DEVICE, std::string("finish: ") + hipGetErrorString(e));
    free_volume(h);  // ownership of the map passes to the caller (reconstructor.cpp:1633-1640)
    return FH_OK;
}

int fh_merge_add(fh_reconstructor *h, const float *threed, const float *normv, int where)
{
    if (!h) return FH_ERR_NULL;
    if (!threed || !normv) return fail(h, FH_ERR_NULL, "NullPointerException: task volume or normout is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "merge_add() before setup()");
    FH_TRY(h, hipSetDevice(h->device));
    const size_t n = vol_elems(h->P);
    const float2 *td = reinterpret_cast<const float2 *>(threed);
    const float *nd = normv;
    void *tmp = nullptr;
    if (where != FH_DEVICE) {
        // staged through one temporary device buffer (12 bytes per stored voxel)
        hipError_t e = hipMalloc(&tmp, (sizeof(float2) + sizeof(float)) * n);
        if (e != hipSuccess) return fail(h, FH_ERR_NOMEM, std::string("merge staging: ") + hipGetErrorString(e));
        float2 *t2 = static_cast<float2 *>(tmp);
        float *n2 = reinterpret_cast<float *>(t2 + n);
        e = hipMemcpyAsync(t2, threed, sizeof(float2) * n, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(n2, normv, sizeof(float) * n, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) {
            (void)hipFree(tmp);
            return fail(h, FH_ERR_DEVICE, std::string("merge staging: ") + hipGetErrorString(e));
        }
        td = t2;
        nd = n2;
    }
    hipError_t e = launch_merge_add(h->data, h->norm, td, nd, n, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) return fail(h, FH_ERR_DEVICE, std::string("merge_add: ") + hipGetErrorString(e));
    return FH_OK;
}

int fh_merge_finish(fh_reconstructor *h, float *out)
{
    if (!h) return FH_ERR_NULL;
    if (!out) return fail(h, FH_ERR_NULL, "NullPointerException: output buffer is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "merge_finish() before setup()");
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P;
    const size_t n = vol_elems(P);
    hipError_t e = launch_merge_scale(h->data, h->norm, n, h->stream);
    if (e == hipSuccess) e = launch_fft3d_inverse_zy(h->data, P, h->plan, h->tw, h->stream);
    float *real = nullptr;
    if (e == hipSuccess) e = hipMalloc(&real, sizeof(float) * (size_t)P * P * P);
    if (e == hipSuccess) {
        const float sc = 1.0f / (float)((size_t)P * P * P);
        e = launch_c2r_x(h->data, P, real, sc, h->plan, h->tw, h->stream);
        if (e == hipSuccess) e = d2h_staged(h, out, real, sizeof(float) * (size_t)P * P * P);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(real);
    }
    if (e != hipSuccess) return fail(h, FH_ERR_DEVICE, std::string("merge_finish: ") + hipGetErrorString(e));
    free_volume(h);
    return FH_OK;
}

// ------------------------------------------------------------ gauss_fft projector
int fh_gauss_fft_volume(fh_reconstructor *h, const float *map)
{
    if (!h) return FH_ERR_NULL;
    if (!map) return fail(h, FH_ERR_NULL, "NullPointerException: volume is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P;
    int rc = ensure_chunk(h, std::max(P, h->cap_n), std::max(P, h->cap_B));
    if (rc) return rc;
    if (!h->pvol) FH_TRY(h, hipMalloc(&h->pvol, sizeof(float2) * vol_elems(P)));
    // xform.phaseorigin.tocorner and do_fft (projector.cpp:77-79) on the GPU; the
    // Fourier origin shift to the centre is a re-indexing, done by the reads
    FH_TRY(h, h2d_staged(h, h->imgs, map, sizeof(float) * (size_t)P * P * P));
    FH_TRY(h, launch_fill_f32(h->means, 0.f, P, h->stream));
    FH_TRY(h, launch_fft3d_forward_corner(h->imgs, P, h->means, h->pvol, h->plan, h->tw, h->stream));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_gauss_fft_project(fh_reconstructor *h, int n, const float *xforms, int mode, int returnfft, float *out)
{
    if (!h) return FH_ERR_NULL;
    if (!xforms || !out) return fail(h, FH_ERR_NULL, "NullPointerException: transforms or output is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->pvol) return fail(h, FH_ERR_STATE, "gauss_fft projection before its volume was set");
    if (mode < 0 || mode > 2)
        return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: gauss_fft modes 0, 1 and 2 are supported");
    if (n < 0) return fail(h, FH_ERR_INVALID_PARAM, "InvalidParameterException: negative projection count");
    FH_TRY(h, hipSetDevice(h->device));
    const int P = h->P, H = h->H;
    // per projection: inverse of the rotation part (get_rotation_transform, invert:
    // transform.cpp:711-719, 1324-1356), the scale and the translation (get_trans)
    std::vector<float> rec(12 * (size_t)std::max(n, 1)), xf6;
    bool any_shift = false;
    for (int q = 0; q < n; q++) {
        const float *t = xforms + 12 * (size_t)q;
        if (determinant(t) < 0)
            return fail(h, FH_ERR_INVALID_PARAM,
                        "InvalidParameterException: gauss_fft with a mirrored Transform is not supported by fourier_hip");
        float r9[9];
        rotation_only(t, r9);
        const double m00 = r9[0], m01 = r9[1], m02 = r9[2], m10 = r9[3], m11 = r9[4], m12 = r9[5];
        const double m20 = r9[6], m21 = r9[7], m22 = r9[8];
        const double cof00 = m11 * m22 - m12 * m21, cof11 = m22 * m00 - m20 * m02, cof22 = m00 * m11 - m01 * m10;
        const double cof01 = m10 * m22 - m20 * m12, cof02 = m10 * m21 - m20 * m11, cof12 = m00 * m21 - m01 * m20;
        const double cof10 = m01 * m22 - m02 * m21, cof20 = m01 * m12 - m02 * m11, cof21 = m00 * m12 - m10 * m02;
        const double det = m00 * cof00 + m02 * cof02 - m01 * cof01;
        float *o = &rec[12 * (size_t)q];
        o[0] = (float)(cof00 / det);
        o[1] = -(float)(cof10 / det);
        o[2] = (float)(cof20 / det);
        o[3] = -(float)(cof01 / det);
        o[4] = (float)(cof11 / det);
        o[5] = -(float)(cof21 / det);
        o[6] = (float)(cof02 / det);
        o[7] = -(float)(cof12 / det);
        o[8] = (float)(cof22 / det);
        o[9] = get_scale(t);
        float tx = t[3], ty = t[7];
        apply_precision(tx, ERR_LIMIT);
        apply_precision(ty, ERR_LIMIT);
        o[10] = tx;
        o[11] = ty;
        any_shift = any_shift || tx != 0.0f || ty != 0.0f;
    }
    const float gw = mode == 2 ? (float)(4.0 / (M_PI * M_PI)) : 1.0f;  // EMConsts::I2G for mode 2
    const int chunk = std::max(1, std::min(n, 4096));
    int rc = ensure_chunk(h, std::max(chunk, h->cap_n), std::max(P, h->cap_B));
    if (rc) return rc;
    if (!returnfft && (rc = ensure_ximgs(h, chunk))) return rc;
    float *tmp = nullptr;  // the projections' records on the device
    FH_TRY(h, hipMalloc(&tmp, sizeof(float) * 12 * (size_t)chunk));
    float *rec_d = tmp;
    hipError_t e = hipSuccess;
    for (int c0 = 0; c0 < n && e == hipSuccess; c0 += chunk) {
        const int m = std::min(chunk, n - c0);
        e = hipMemcpyAsync(rec_d, &rec[12 * (size_t)c0], sizeof(float) * 12 * m, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = launch_gauss_fft(h->pvol, P, m, mode, gw, rec_d, returnfft ? 1 : 0, h->slices, h->stream);
        if (e != hipSuccess) break;
        if (returnfft) {
            e = hipMemcpyAsync(out + (size_t)c0 * P * (P + 2), h->slices, sizeof(float2) * (size_t)m * P * (H + 1),
                               hipMemcpyDeviceToHost, h->stream);
        } else {
            // do_ift (x 1/P^2) written with the phase origin moved to the centre
            e = launch_ift2d_batch(h->slices, m, P, h->ximgs, 1.0f / ((float)P * (float)P), h->plan, h->tw, h->stream);
            float *img = h->ximgs;
            if (e == hipSuccess && any_shift) {
                // translate(): the "xform" processor's real-space resampling by the translation
                xf6.assign(6 * (size_t)m, 0.f);
                for (int q = 0; q < m; q++) {
                    const float *o = &rec[12 * (size_t)(c0 + q)];
                    float tr[12] = {1, 0, 0, o[10], 0, 1, 0, o[11], 0, 0, 1, 0};
                    float inv[6] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
                    xform2d_inverse(tr, inv);
                    memcpy(&xf6[6 * (size_t)q], inv, sizeof(inv));
                }
                if ((rc = ensure_ximgs(h, chunk, true))) {
                    (void)hipFree(tmp);
                    return rc;
                }
                e = hipMemcpyAsync(h->xfd, xf6.data(), sizeof(float) * 6 * m, hipMemcpyHostToDevice, h->stream);
                if (e == hipSuccess) e = launch_fill_f32(h->means, 0.f, m, h->stream);
                if (e == hipSuccess) e = launch_xform2d(h->ximgs, m, P, P, h->means, h->xfd, h->ximgs2, h->stream);
                img = h->ximgs2;
            }
            if (e == hipSuccess) e = launch_gausskernelfix2d(img, P, m, gw, h->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(out + (size_t)c0 * P * P, img, sizeof(float) * (size_t)m * P * P, hipMemcpyDeviceToHost,
                                   h->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(h, FH_ERR_DEVICE, std::string("gauss_fft: ") + hipGetErrorString(e));
    return FH_OK;
}

int fh_get_state(fh_reconstructor *h, float *data, float *weights)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->ready) return fail(h, FH_ERR_STATE, "get_state() before setup()");
    FH_TRY(h, hipSetDevice(h->device));
    const size_t n = vol_elems(h->P);
    if (data) FH_TRY(h, d2h_staged(h, data, h->data, sizeof(float2) * n));
    if (weights) FH_TRY(h, d2h_staged(h, weights, h->norm, sizeof(float) * n));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_set_tuning(fh_reconstructor *h, const fh_tuning *t)
{
    if (!h || !t) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (t->waves_per_block < 1 || t->waves_per_block > 4)
        return fail(h, FH_ERR_INVALID_PARAM, "waves_per_block must be in [1, 4]");
    if (t->slices_per_item < 1 || t->chunk < 0 || t->blocks_per_cu < 1)
        return fail(h, FH_ERR_INVALID_PARAM, "tuning values must be positive (chunk 0: automatic)");
    if (t->graph != 0 && t->graph != 1) return fail(h, FH_ERR_INVALID_PARAM, "graph must be 0 or 1");
    if (t->slices_per_item != h->tune.slices_per_item) free_bins(h);
    h->tune = *t;
    return FH_OK;
}

int fh_get_tuning(const fh_reconstructor *h, fh_tuning *t)
{
    if (!h || !t) return FH_ERR_NULL;
    *t = h->tune;
    return FH_OK;
}

int fh_get_stats(const fh_reconstructor *h, fh_stats *s)
{
    if (!h || !s) return FH_ERR_NULL;
    *s = h->stats;
    return FH_OK;
}

int fh_sync(fh_reconstructor *h)
{
    if (!h) return FH_ERR_NULL;
    std::lock_guard<std::mutex> lk(h->mu);
    FH_TRY(h, hipSetDevice(h->device));
    FH_TRY(h, hipStreamSynchronize(h->stream));
    return FH_OK;
}

int fh_get_symmetries(const char *sym, float *out9, int max_ops)
{
    std::vector<float> ops;
    if (!sym || !sym_ops(sym, ops)) return -FH_ERR_NOT_EXISTING;
    const int n = (int)ops.size() / 9;
    if (out9) {
        if (max_ops < n) return -FH_ERR_INVALID_PARAM;
        memcpy(out9, ops.data(), sizeof(float) * ops.size());
    }
    return n;
}

long long fh_footprint_bytes(int P, int nsym, int chunk)
{
    return footprint_bytes_for_mode(P, nsym, chunk, nullptr);
}

}  // extern "C"

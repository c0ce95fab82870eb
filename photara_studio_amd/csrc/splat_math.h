// splat_math.h — per-element math of the 3DGS training step, shared by the gfx950 kernels
// (T = float) and the CPU oracle (T = float for bit-exact parity, T = double for accuracy).
//
// Determinism contract: every float expression here is evaluated with IEEE-754 single
// precision, correctly rounded +,-,*,/,sqrt, explicit fmaf where a fused op is wanted, and
// no contraction (compile with -ffp-contract=off, no fast-math). exp/log are this file's
// own routines (pexpf/plogf) built only from those operations, so the CPU and the GPU
// produce identical bits for keys, tile ranges, contributor counts and images.
//
// Formulas restate the reference rasterizer (third_party/splat_drender/src/device/*.cuh,
// src/render_forward.cu, src/render_backward.cu); each function cites what it follows.
#pragma once

#include <math.h>
#include <stdint.h>

#include "pht_splat.h"

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define PHT_HD __host__ __device__ inline
#else
#define PHT_HD inline
#endif

namespace pht {

// ---------------------------------------------------------------- constants
// third_party/splat_drender/include/splat_drender/config.h:35-41
template <class T> struct K {
    static constexpr T alpha_floor() { return T(1) / T(255); }
    static constexpr T transmittance_floor() { return T(1.0e-4); }
};

// ---------------------------------------------------------------- scalar helpers
PHT_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
PHT_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

PHT_HD float fma_(float a, float b, float c) { return fmaf(a, b, c); }
PHT_HD double fma_(double a, double b, double c) { return fma(a, b, c); }
PHT_HD float sqrt_(float x) { return sqrtf(x); }
PHT_HD double sqrt_(double x) { return sqrt(x); }
PHT_HD float floor_(float x) { return floorf(x); }
PHT_HD double floor_(double x) { return floor(x); }
PHT_HD float ceil_(float x) { return ceilf(x); }
PHT_HD double ceil_(double x) { return ceil(x); }
PHT_HD float min_(float a, float b) { return fminf(a, b); }
PHT_HD double min_(double a, double b) { return fmin(a, b); }
PHT_HD float max_(float a, float b) { return fmaxf(a, b); }
PHT_HD double max_(double a, double b) { return fmax(a, b); }
PHT_HD float abs_(float a) { return fabsf(a); }
PHT_HD double abs_(double a) { return fabs(a); }
PHT_HD bool finite_(float a) { return (f2u(a) & 0x7f800000u) != 0x7f800000u; }
PHT_HD bool finite_(double a) { return a - a == 0.0; }

// exp(x), Cody-Waite range reduction x = k*ln2 + r (|r| <= ln2/2) with ln2 split into a
// 24-bit head and its remainder, then the degree-7 Taylor polynomial of e^r in Horner form
// (truncation error r^8/8! < 6e-9), then an exact scaling by 2^k. Max error ~1 ulp.
PHT_HD float pexpf(float x) {
    if (!(x > -104.0f)) return x != x ? x : 0.0f;
    if (x > 88.72283935546875f) return u2f(0x7f800000u);
    const float kf = rintf(x * 1.44269502162933350f);
    float r = fma_(kf, -0.693147182464599609375f, x);
    r = fma_(kf, 1.9046542121259336e-09f, r);
    float p = 1.98412701138295233e-04f;
    p = fma_(p, r, 1.38888892251998186e-03f);
    p = fma_(p, r, 8.33333376795053482e-03f);
    p = fma_(p, r, 4.16666679084300995e-02f);
    p = fma_(p, r, 1.66666671633720398e-01f);
    p = fma_(p, r, 0.5f);
    p = fma_(p, r, 1.0f);
    p = fma_(p, r, 1.0f);
    const int k = int(kf);
    if (k > 127) return (p * u2f(uint32_t(127 + 127) << 23)) * 2.0f;
    if (k >= -125) return p * u2f(uint32_t(k + 127) << 23);
    return (p * u2f(uint32_t(k + 127 + 64) << 23)) * 5.42101086242752217e-20f; // 2^-64
}

// log(x): x = 2^e * m with m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s), s = (m-1)/(m+1),
// atanh series to s^9 (|s| <= 0.1716, truncation < 1e-9 relative). ~1-2 ulp.
PHT_HD float plogf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -u2f(0x7f800000u) : u2f(0x7fc00000u);
    if (x == u2f(0x7f800000u)) return x;
    uint32_t b = f2u(x);
    int e = 0;
    if (b < 0x00800000u) { x = x * 8388608.0f; b = f2u(x); e = -23; }
    e += int(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986084f) { m = m * 0.5f; e += 1; }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float s2 = s * s;
    float q = 1.11111111938953400e-01f;
    q = fma_(q, s2, 1.42857149243354797e-01f);
    q = fma_(q, s2, 2.00000002980232239e-01f);
    q = fma_(q, s2, 3.33333343267440796e-01f);
    const float s_2 = 2.0f * s;
    const float lm = fma_(s_2, s2 * q, s_2);
    const float ef = float(e);
    return fma_(ef, 0.693147182464599609375f, fma_(ef, -1.9046542121259336e-09f, lm));
}

PHT_HD float exp_(float x) { return pexpf(x); }
PHT_HD double exp_(double x) { return exp(x); }
PHT_HD float log_(float x) { return plogf(x); }
PHT_HD double log_(double x) { return log(x); }

// ---------------------------------------------------------------- 3x3 (row-major)
// third_party/splat_drender/src/device/matrix.cuh:62-193
template <class T> struct M3 {
    T m[3][3];
    PHT_HD T at(int c, int r) const { return m[r][c]; } // column-major accessor
};
template <class T> PHT_HD M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return o;
}
template <class T> PHT_HD M3<T> transpose(const M3<T>& a) {
    M3<T> o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.m[i][j] = a.m[j][i];
    return o;
}
// Scalar-first quaternion -> rows (matrix.cuh:171-184).
template <class T> PHT_HD M3<T> quat_rotation(const T q[4]) {
    const T r = q[0], x = q[1], y = q[2], z = q[3];
    M3<T> o;
    o.m[0][0] = T(1) - T(2) * (y * y + z * z);
    o.m[0][1] = T(2) * (x * y + r * z);
    o.m[0][2] = T(2) * (x * z - r * y);
    o.m[1][0] = T(2) * (x * y - r * z);
    o.m[1][1] = T(1) - T(2) * (x * x + z * z);
    o.m[1][2] = T(2) * (y * z + r * x);
    o.m[2][0] = T(2) * (x * z + r * y);
    o.m[2][1] = T(2) * (y * z - r * x);
    o.m[2][2] = T(1) - T(2) * (x * x + y * y);
    return o;
}
// Column-major world_to_camera applied to a point (matrix.cuh:142-146).
template <class T> PHT_HD void xform_point(const T p[3], const float* v, T t[3]) {
    t[0] = ((T(v[0]) * p[0] + T(v[4]) * p[1]) + T(v[8]) * p[2]) + T(v[12]);
    t[1] = ((T(v[1]) * p[0] + T(v[5]) * p[1]) + T(v[9]) * p[2]) + T(v[13]);
    t[2] = ((T(v[2]) * p[0] + T(v[6]) * p[1]) + T(v[10]) * p[2]) + T(v[14]);
}
template <class T> PHT_HD void xform_dir_transpose(const T d[3], const float* v, T o[3]) {
    o[0] = (T(v[0]) * d[0] + T(v[1]) * d[1]) + T(v[2]) * d[2];
    o[1] = (T(v[4]) * d[0] + T(v[5]) * d[1]) + T(v[6]) * d[2];
    o[2] = (T(v[8]) * d[0] + T(v[9]) * d[1]) + T(v[10]) * d[2];
}
template <class T> PHT_HD M3<T> world_rotation_transposed(const float* v) {
    M3<T> w;
    w.m[0][0] = T(v[0]); w.m[0][1] = T(v[1]); w.m[0][2] = T(v[2]);
    w.m[1][0] = T(v[4]); w.m[1][1] = T(v[5]); w.m[1][2] = T(v[6]);
    w.m[2][0] = T(v[8]); w.m[2][1] = T(v[9]); w.m[2][2] = T(v[10]);
    return w;
}

// ---------------------------------------------------------------- activation
// activate_row (third_party/splat_drender/src/device/activation.cuh:13-31) ==
// activate_kernel (photara/src/splat/cuda_ops.cu:208-239).
template <class T> struct Activated {
    T scale[3];     // filtered scale sqrt(exp(ls)^2 + f^2)
    T raw_scale[3]; // exp(ls)
    T rot[4];       // normalised quaternion
    T inv_norm;     // 1 / max(|q|, 1e-10)
    T sigmoid;      // sigmoid(logit)
    T opacity;      // sigmoid * prod(raw/filtered)
};
template <class T>
PHT_HD Activated<T> activate(const T ls[3], const T q[4], T logit, bool has_filter, T filter) {
    Activated<T> a;
    T ratio = T(1);
    const T f2 = has_filter ? filter * filter : T(0);
    for (int k = 0; k < 3; ++k) {
        const T v = exp_(ls[k]);
        a.raw_scale[k] = v;
        if (has_filter) {
            a.scale[k] = sqrt_(v * v + f2);
            ratio = ratio * (v / a.scale[k]);
        } else {
            a.scale[k] = v;
        }
    }
    const T n2 = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    const T inv = T(1) / sqrt_(max_(n2, T(1e-20)));
    a.inv_norm = inv;
    for (int k = 0; k < 4; ++k) a.rot[k] = q[k] * inv;
    a.sigmoid = T(1) / (T(1) + exp_(-logit));
    a.opacity = a.sigmoid * ratio;
    return a;
}

// ---------------------------------------------------------------- camera (internal)
// Per-view constants as kernels see them: pht_camera plus the derived tile grid and the
// view's offsets into the global tile / pixel arrays of a batched launch.
struct ViewInfo {
    pht_camera cam;
    int32_t grid_x, grid_y;
    int32_t tile_base;  // first global tile id of this view
    int32_t tile_count;
    int64_t pixel_base; // first global pixel of this view
};

// ---------------------------------------------------------------- screen footprint cap
// footprint::policy / apply / conic_backward (third_party/splat_drender/src/device/footprint.cuh:13-48).
// The eigen-radius uses sqrt(h^2 + b^2) where the reference calls hypotf.
template <class T> struct Cap { T q, lambda, da, db, dc; };
template <class T> PHT_HD Cap<T> cap_policy(T a, T b, T c, int width, int height) {
    const T half = T(0.5) * (a - c);
    const T radius = sqrt_(half * half + b * b);
    const T lambda = (T(0.5) * a + T(0.5) * c) + radius;
    const int mn = width < height ? width : height;
    const T limit = min_(T(1024), T(mn));
    const T variance = (T(0.125) * limit) * limit;
    Cap<T> cp;
    cp.lambda = lambda;
    if (!(lambda > variance) || !finite_(lambda)) {
        cp.q = T(1); cp.da = T(0); cp.db = T(0); cp.dc = T(0);
        return cp;
    }
    const T h = radius > T(0) ? half / radius : T(0);
    cp.q = variance / lambda;
    cp.da = T(0.5) + T(0.5) * h;
    cp.db = radius > T(0) ? b / radius : T(0);
    cp.dc = T(0.5) - T(0.5) * h;
    return cp;
}
template <class T>
PHT_HD void conic_backward_capped(T a, T b, T c, const Cap<T>& cap, const T dconic[3], T out[3]) {
    const T x = a * cap.q, y = b * cap.q, z = c * cap.q;
    const T det = x * z - y * y;
    const T u = z / det, v = -y / det, w = x / det;
    const T ga = -((u * u * dconic[0] + T(2) * u * v * dconic[1]) + v * v * dconic[2]);
    const T gb = T(-2) * ((u * v * dconic[0] + (u * w + v * v) * dconic[1]) + v * w * dconic[2]);
    const T gc = -((v * v * dconic[0] + T(2) * v * w * dconic[1]) + w * w * dconic[2]);
    const T pull = ((ga * a + gb * b) + gc * c) / cap.lambda;
    out[0] = cap.q * (ga - pull * cap.da);
    out[1] = cap.q * (gb - pull * cap.db);
    out[2] = cap.q * (gc - pull * cap.dc);
}

// ---------------------------------------------------------------- EWA projection (pinhole)
// geo::project_splat pinhole branch (geometry.cuh:127-163,178-198,248-253) fused with the
// conic / radius part of preprocess_gaussians (render_forward.cu:69-83).
// Screen covariance: M = S R (W J), cov = M^T M (only the 2x2 block is live since J's third
// column is zero), + kernel_size on the diagonal, capped, then inverted.
template <class T> struct Projected {
    T mean2d[2];
    T conic[3];
    T opacity;   // activated opacity * compensation
    T depth;     // |t| (radial distance)
    int radius;
    bool fp64;   // 2x2 inverse went through the FP64 route
};

// Relative determinant below which the float 2x2 inverse loses more than 8 bits to
// cancellation; such splats are inverted in FP64 when fp64_conic is set.
constexpr float kNearSingular = 1.0f / 256.0f;
// Relative determinant below which the backward of the 2x2 inverse (conic -> covariance) is
// evaluated in FP64 when fp64_conic is set. Its terms cancel by about a c / det, and the scale
// gradient along an elongated splat's long axis cancels again when it is contracted with the
// projection; in float the per-view gradient of such a splat's long log-scale loses up to 1e-3
// relative (anisotropy ~150 in 2D). The FP64 evaluation takes the same float inputs and rounds
// its three results to float.
constexpr float kIllConditioned = 1.0f / 4.0f;

// The conic backward (geometry.cuh:759-809 uncapped branch: dL/d(a, b, c) of the 2D covariance
// from dL/d(conic)) in double precision, from and to float.
template <class T>
PHT_HD void conic_backward_f64(T a, T b, T c, const T dconic[3], T& ga, T& gb, T& gc) {
    const double A = double(a), B = double(b), Cv = double(c);
    const double d0 = double(dconic[0]), d1 = double(dconic[1]), d2 = double(dconic[2]);
    const double denom = A * Cv - B * B;
    const double inv = 1.0 / (denom * denom + 1.0e-7);
    ga = T(inv * ((-Cv * Cv * d0 + 2.0 * B * Cv * d1) + (denom - A * Cv) * d2));
    gc = T(inv * ((-A * A * d2 + 2.0 * A * B * d1) + (denom - A * Cv) * d0));
    gb = T(inv * 2.0 * ((B * Cv * d0 - (denom + 2.0 * B * B) * d1) + A * B * d2));
}

template <class T>
PHT_HD bool project_pinhole(const T mean[3], const Activated<T>& act, const pht_camera& cam,
                            const pht_settings& s, Projected<T>& out) {
    T t[3];
    xform_point(mean, cam.world_to_camera, t);
    if (!(t[2] > T(s.near_plane))) return false; // camera.h:241-246
    const T fx = T(cam.fx), fy = T(cam.fy);
    const T limx = T(1.3) * (T(cam.width) / (T(2) * fx));
    const T limy = T(1.3) * (T(cam.height) / (T(2) * fy));
    const T u = t[0] / t[2], v = t[1] / t[2];
    const T txc = min_(limx, max_(-limx, u)) * t[2];
    const T tyc = min_(limy, max_(-limy, v)) * t[2];
    const T tz2 = t[2] * t[2];
    const T j00 = fx / t[2], j11 = fy / t[2];
    const T j20 = -(fx * txc) / tz2, j21 = -(fy * tyc) / tz2;
    // T = W J: row i = (W[i][0] J00 + W[i][2] J20, W[i][1] J11 + W[i][2] J21, 0)
    const M3<T> W = world_rotation_transposed<T>(cam.world_to_camera);
    T tm[3][2];
    for (int i = 0; i < 3; ++i) {
        tm[i][0] = W.m[i][0] * j00 + W.m[i][2] * j20;
        tm[i][1] = W.m[i][1] * j11 + W.m[i][2] * j21;
    }
    const M3<T> R = quat_rotation<T>(act.rot);
    const T sm = T(s.scale_modifier);
    T mm[3][2]; // M = (S R) T
    for (int i = 0; i < 3; ++i) {
        const T si = sm * act.scale[i];
        const T r0 = si * R.m[i][0], r1 = si * R.m[i][1], r2 = si * R.m[i][2];
        mm[i][0] = (r0 * tm[0][0] + r1 * tm[1][0]) + r2 * tm[2][0];
        mm[i][1] = (r0 * tm[0][1] + r1 * tm[1][1]) + r2 * tm[2][1];
    }
    const T c00 = (mm[0][0] * mm[0][0] + mm[1][0] * mm[1][0]) + mm[2][0] * mm[2][0];
    const T c01 = (mm[0][0] * mm[0][1] + mm[1][0] * mm[1][1]) + mm[2][0] * mm[2][1];
    const T c11 = (mm[0][1] * mm[0][1] + mm[1][1] * mm[1][1]) + mm[2][1] * mm[2][1];
    const T ks = T(s.kernel_size);
    T a = c00 + ks, b = c01, c = c11 + ks;
    T coef = T(1);
    if (s.compensate_2d) {
        const T det0 = max_(T(1.0e-6), c00 * c11 - c01 * c01);
        const T det1 = max_(T(1.0e-6), a * c - b * b);
        coef = sqrt_(det0 / det1);
    }
    // footprint::apply (footprint.cuh:30-37)
    const Cap<T> cp = cap_policy(a, b, c, cam.width, cam.height);
    if (cp.q < T(1)) { a = a * cp.q; b = b * cp.q; c = c * cp.q; }
    // conic (render_forward.cu:74-79)
    T ac = a * c;
    T det = ac - b * b;
    out.fp64 = false;
    if (s.fp64_conic && sizeof(T) == 4 && !(det > T(kNearSingular) * ac)) {
        // Near-singular 2x2 block: the float products cancel. Form the 2x2 covariance, its cap,
        // determinant and inverse in FP64 from the same float inputs, then round back.
        double md[3][2];
        for (int i = 0; i < 3; ++i) {
            const double si = double(sm) * double(act.scale[i]);
            const double r0 = si * double(R.m[i][0]), r1 = si * double(R.m[i][1]), r2 = si * double(R.m[i][2]);
            md[i][0] = (r0 * double(tm[0][0]) + r1 * double(tm[1][0])) + r2 * double(tm[2][0]);
            md[i][1] = (r0 * double(tm[0][1]) + r1 * double(tm[1][1])) + r2 * double(tm[2][1]);
        }
        double da = ((md[0][0] * md[0][0] + md[1][0] * md[1][0]) + md[2][0] * md[2][0]) + double(ks);
        double db = (md[0][0] * md[0][1] + md[1][0] * md[1][1]) + md[2][0] * md[2][1];
        double dc = ((md[0][1] * md[0][1] + md[1][1] * md[1][1]) + md[2][1] * md[2][1]) + double(ks);
        const Cap<double> cpd = cap_policy(da, db, dc, cam.width, cam.height);
        if (cpd.q < 1.0) { da = da * cpd.q; db = db * cpd.q; dc = dc * cpd.q; }
        const double dd = da * dc - db * db;
        if (!(dd > 0.0)) return false;
        const double di = 1.0 / dd;
        out.conic[0] = T(dc * di);
        out.conic[1] = T(-db * di);
        out.conic[2] = T(da * di);
        a = T(da); b = T(db); c = T(dc);
        det = T(dd);
        out.fp64 = true;
    } else {
        if (det == T(0)) return false;
        const T det_inv = T(1) / det;
        out.conic[0] = c * det_inv;
        out.conic[1] = -b * det_inv;
        out.conic[2] = a * det_inv;
    }
    out.opacity = act.opacity * coef;
    // radius (render_forward.cu:81-83)
    const T mid = T(0.5) * (a + c);
    const T root = sqrt_(max_(T(0.1), mid * mid - det));
    out.radius = int(ceil_(T(3) * sqrt_(max_(mid + root, mid - root))));
    // pixel centre from the unclamped point (camera.h:101-116)
    const T iz = T(1) / t[2];
    out.mean2d[0] = (t[0] * iz) * fx + T(cam.cx);
    out.mean2d[1] = (t[1] * iz) * fy + T(cam.cy);
    out.depth = sqrt_((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]); // render_forward.cu:104
    return true;
}

// ---------------------------------------------------------------- tile enumeration
// Opacity-aware AccuTile footprint (third_party/splat_drender/src/device/tiles.cuh:18-179,
// non-wrapping path): the ellipse where alpha falls to 1/255, walked along the shorter screen
// axis with an exact per-slice interval. Split into the walk's set-up (TileWalk), one column's
// tile interval (tile_walk_column) and the full walk (enumerate_tiles, emit(tx, ty) in walk
// order), all with the same arithmetic.
template <class T> struct TileWalk {
    bool ok;
    bool iterate_x;   // columns run along x (u = tile x, v = tile y); else u = y, v = x
    int gu, gv, u_min, u_max;
    T cu, cv, b, det, threshold, eu, ev, arg_v_min, arg_v_max, tsf, inv_ts;
    T av, inv_av;
};

// The tile edge is 8 or 16, so x / ts and x * (1 / ts) are the same number; the per-column
// quotients by av use one reciprocal per splat (the oracle runs this same code).
template <class T>
PHT_HD TileWalk<T> tile_walk_setup(T mx, T my, const T conic[3], T op, int grid_x, int grid_y, int ts) {
    TileWalk<T> w;
    w.ok = false;
    const T a = conic[0], b = conic[1], c = conic[2];
    const T det = a * c - b * b;
    if (!(a > T(0) && c > T(0)) || !(op >= K<T>::alpha_floor()) || !(det > T(0))) return w;
    const T threshold = T(2) * log_(op * T(255));
    if (!(threshold >= T(0))) return w;
    const T ex = sqrt_(threshold * c / det);
    const T ey = sqrt_(threshold * a / det);
    const T inv_ts = T(1) / T(ts);
    const bool iterate_x = ex <= ey; // walk the shorter screen extent
    w.iterate_x = iterate_x;
    w.cu = iterate_x ? mx : my;
    w.cv = iterate_x ? my : mx;
    const T au = iterate_x ? a : c;
    w.av = iterate_x ? c : a;
    w.inv_av = T(1) / w.av;
    w.gu = iterate_x ? grid_x : grid_y;
    w.gv = iterate_x ? grid_y : grid_x;
    w.b = b;
    w.det = det;
    w.threshold = threshold;
    w.tsf = T(ts);
    w.inv_ts = inv_ts;
    // the extents along u and v are ex and ey in walk order (the same operations: av is c when
    // walking x, so threshold * av / det is threshold * c / det)
    w.eu = iterate_x ? ex : ey;
    w.ev = iterate_x ? ey : ex;
    int u_min = int(floor_((w.cu - w.eu) * inv_ts));
    u_min = u_min < w.gu ? u_min : w.gu;
    u_min = u_min > 0 ? u_min : 0;
    int u_max = int(floor_((w.cu + w.eu) * inv_ts)) + 1;
    u_max = u_max < w.gu ? u_max : w.gu;
    w.u_min = u_min;
    w.u_max = u_max;
    if (u_min >= u_max) return w;
    w.arg_v_min = w.cu + b * w.ev / au;
    w.arg_v_max = w.cu - b * w.ev / au;
    w.ok = true;
    return w;
}

// Tile interval [v_min, v_max) of column tu (may be empty).
template <class T> PHT_HD void tile_walk_column(const TileWalk<T>& w, int tu, int& v_min, int& v_max) {
    const T cu = w.cu, cv = w.cv, av = w.av, b = w.b, det = w.det;
    const T s0 = max_(cu - w.eu, T(tu) * w.tsf);
    const T s1 = min_(cu + w.eu, T(tu + 1) * w.tsf);
    const T d0 = s0 - cu, d1 = s1 - cu;
    const T r0 = sqrt_(max_(T(0), av * w.threshold - det * d0 * d0));
    const T r1 = sqrt_(max_(T(0), av * w.threshold - det * d1 * d1));
    const T lo0 = cv + (-b * d0 - r0) * w.inv_av;
    const T lo1 = cv + (-b * d1 - r1) * w.inv_av;
    const T hi0 = cv + (-b * d0 + r0) * w.inv_av;
    const T hi1 = cv + (-b * d1 + r1) * w.inv_av;
    const T sv_min = (w.arg_v_min >= s0 && w.arg_v_min <= s1) ? cv - w.ev : min_(lo0, lo1);
    const T sv_max = (w.arg_v_max >= s0 && w.arg_v_max <= s1) ? cv + w.ev : max_(hi0, hi1);
    int lo = int(floor_(sv_min * w.inv_ts));
    lo = lo < w.gv ? lo : w.gv;
    lo = lo > 0 ? lo : 0;
    int hi = int(floor_(sv_max * w.inv_ts)) + 1;
    hi = hi < w.gv ? hi : w.gv;
    v_min = lo;
    v_max = hi;
}

template <class T, class Emit>
PHT_HD uint32_t enumerate_tiles(T mx, T my, const T conic[3], T op, int grid_x, int grid_y, int ts,
                                Emit&& emit) {
    const TileWalk<T> w = tile_walk_setup(mx, my, conic, op, grid_x, grid_y, ts);
    if (!w.ok) return 0;
    uint32_t count = 0;
#pragma unroll 1
    for (int tu = w.u_min; tu < w.u_max; ++tu) {
        int v_min, v_max;
        tile_walk_column(w, tu, v_min, v_max);
#pragma unroll 1
        for (int tv = v_min; tv < v_max; ++tv) {
            emit(w.iterate_x ? tu : tv, w.iterate_x ? tv : tu);
            ++count;
        }
    }
    return count;
}

// ---------------------------------------------------------------- spherical harmonics
// Real SH up to degree 3 (third_party/splat_drender/src/device/sh.cuh:8-65): the standard
// normalisation constants, rgb = max(sum_k Y_k c_k + 0.5, 0) per channel.
constexpr double kShC0 = 0.28209479177387814;
constexpr double kShC1 = 0.4886025119029199;
constexpr double kShC2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                             -1.0925484305920792, 0.5462742152960396};
constexpr double kShC3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                             0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                             -0.5900435899266435};

PHT_HD int sh_count(int degree) { return (degree + 1) * (degree + 1); }

// Unit view direction from the camera centre (sh.cuh:20-23).
template <class T> PHT_HD void view_dir(const T mean[3], const float pos[3], T d[3], T raw[3]) {
    raw[0] = mean[0] - T(pos[0]);
    raw[1] = mean[1] - T(pos[1]);
    raw[2] = mean[2] - T(pos[2]);
    const T inv = T(1) / sqrt_((raw[0] * raw[0] + raw[1] * raw[1]) + raw[2] * raw[2]);
    d[0] = raw[0] * inv; d[1] = raw[1] * inv; d[2] = raw[2] * inv;
}

template <class T> PHT_HD void sh_basis(int degree, const T d[3], T Y[16]) {
    const T x = d[0], y = d[1], z = d[2];
    Y[0] = T(kShC0);
    if (degree < 1) return;
    Y[1] = -T(kShC1) * y;
    Y[2] = T(kShC1) * z;
    Y[3] = -T(kShC1) * x;
    if (degree < 2) return;
    const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    Y[4] = T(kShC2[0]) * xy;
    Y[5] = T(kShC2[1]) * yz;
    Y[6] = T(kShC2[2]) * ((T(2) * zz - xx) - yy);
    Y[7] = T(kShC2[3]) * xz;
    Y[8] = T(kShC2[4]) * (xx - yy);
    if (degree < 3) return;
    Y[9] = T(kShC3[0]) * (y * (T(3) * xx - yy));
    Y[10] = T(kShC3[1]) * (xy * z);
    Y[11] = T(kShC3[2]) * (y * ((T(4) * zz - xx) - yy));
    Y[12] = T(kShC3[3]) * (z * ((T(2) * zz - T(3) * xx) - T(3) * yy));
    Y[13] = T(kShC3[4]) * (x * ((T(4) * zz - xx) - yy));
    Y[14] = T(kShC3[5]) * (z * (xx - yy));
    Y[15] = T(kShC3[6]) * (x * (xx - T(3) * yy));
}

// coef[k*3 + ch]; writes rgb and the per-channel clamp flags (bit ch set when clamped).
template <class T>
PHT_HD void sh_eval(int degree, const T d[3], const T* coef, T rgb[3], uint32_t& clamp_bits) {
    T Y[16];
    sh_basis(degree, d, Y);
    const int nb = sh_count(degree);
    clamp_bits = 0;
    for (int ch = 0; ch < 3; ++ch) {
        T acc = Y[0] * coef[ch];
        for (int k = 1; k < nb; ++k) acc = fma_(Y[k], coef[k * 3 + ch], acc);
        acc = acc + T(0.5);
        if (acc < T(0)) clamp_bits |= 1u << ch;
        rgb[ch] = max_(acc, T(0));
    }
}

// d(unit dir)/d(raw dir) applied to dv (matrix.cuh:187-193).
template <class T> PHT_HD void dnormalize(const T v[3], const T dv[3], T o[3]) {
    const T inv = T(1) / sqrt_((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const T n[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    o[0] = (((T(1) - n[0] * n[0]) * dv[0] - n[1] * n[0] * dv[1]) - n[2] * n[0] * dv[2]) * inv;
    o[1] = ((-n[0] * n[1] * dv[0] + (T(1) - n[1] * n[1]) * dv[1]) - n[2] * n[1] * dv[2]) * inv;
    o[2] = ((-n[0] * n[2] * dv[0] - n[1] * n[2] * dv[1]) + (T(1) - n[2] * n[2]) * dv[2]) * inv;
}

// Backward of sh_eval (sh.cuh:75-171), in two parts that can run in separate passes:
// sh_backward_coef accumulates dL/dcoef into gsh[k*3+ch]; sh_backward_dir adds the view-direction
// term to gmean (through the normalisation of mean - camera). Clamped channels carry no gradient.
template <class T>
PHT_HD void sh_backward_coef(int degree, const T d[3], uint32_t clamp_bits, const T dl_rgb[3], T* gsh) {
    T dl[3];
    for (int ch = 0; ch < 3; ++ch) dl[ch] = (clamp_bits >> ch) & 1u ? T(0) : dl_rgb[ch];
    T Y[16];
    sh_basis(degree, d, Y);
    const int nb = sh_count(degree);
    for (int k = 0; k < nb; ++k)
        for (int ch = 0; ch < 3; ++ch) gsh[k * 3 + ch] += Y[k] * dl[ch];
}

// coef: the coefficients, anything indexable as coef[k * 3 + ch] (a pointer, or a device-side view
// of a packed reduced-precision row).
template <class T, class C>
PHT_HD void sh_backward_dir(int degree, const T raw_dir[3], const T d[3], const C& coef, uint32_t clamp_bits,
                            const T dl_rgb[3], T gmean[3]) {
    if (degree < 1) return;
    T dl[3];
    for (int ch = 0; ch < 3; ++ch) dl[ch] = (clamp_bits >> ch) & 1u ? T(0) : dl_rgb[ch];
    const int nb = sh_count(degree);
    const T x = d[0], y = d[1], z = d[2];
    // dY_k/d(x,y,z) for k >= 1
    T dYx[16], dYy[16], dYz[16];
    dYx[1] = T(0); dYy[1] = -T(kShC1); dYz[1] = T(0);
    dYx[2] = T(0); dYy[2] = T(0); dYz[2] = T(kShC1);
    dYx[3] = -T(kShC1); dYy[3] = T(0); dYz[3] = T(0);
    if (degree > 1) {
        const T c0 = T(kShC2[0]), c1 = T(kShC2[1]), c2 = T(kShC2[2]), c3 = T(kShC2[3]), c4 = T(kShC2[4]);
        dYx[4] = c0 * y;            dYy[4] = c0 * x;            dYz[4] = T(0);
        dYx[5] = T(0);              dYy[5] = c1 * z;            dYz[5] = c1 * y;
        dYx[6] = c2 * (T(-2) * x);  dYy[6] = c2 * (T(-2) * y);  dYz[6] = c2 * (T(4) * z);
        dYx[7] = c3 * z;            dYy[7] = T(0);              dYz[7] = c3 * x;
        dYx[8] = c4 * (T(2) * x);   dYy[8] = c4 * (T(-2) * y);  dYz[8] = T(0);
    }
    if (degree > 2) {
        const T xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
        const T k0 = T(kShC3[0]), k1 = T(kShC3[1]), k2 = T(kShC3[2]), k3 = T(kShC3[3]);
        const T k4 = T(kShC3[4]), k5 = T(kShC3[5]), k6 = T(kShC3[6]);
        dYx[9] = k0 * (T(6) * xy);                    dYy[9] = k0 * (T(3) * (xx - yy));
        dYz[9] = T(0);
        dYx[10] = k1 * yz;                            dYy[10] = k1 * xz;
        dYz[10] = k1 * xy;
        dYx[11] = k2 * (T(-2) * xy);                  dYy[11] = k2 * ((T(-3) * yy + T(4) * zz) - xx);
        dYz[11] = k2 * (T(8) * yz);
        dYx[12] = k3 * (T(-6) * xz);                  dYy[12] = k3 * (T(-6) * yz);
        dYz[12] = k3 * (T(3) * ((T(2) * zz - xx) - yy));
        dYx[13] = k4 * ((T(-3) * xx + T(4) * zz) - yy); dYy[13] = k4 * (T(-2) * xy);
        dYz[13] = k4 * (T(8) * xz);
        dYx[14] = k5 * (T(2) * xz);                   dYy[14] = k5 * (T(-2) * yz);
        dYz[14] = k5 * (xx - yy);
        dYx[15] = k6 * (T(3) * (xx - yy));            dYy[15] = k6 * (T(-6) * xy);
        dYz[15] = T(0);
    }
    T dd[3] = {T(0), T(0), T(0)};
    for (int ch = 0; ch < 3; ++ch) {
        T sx = T(0), sy = T(0), sz = T(0);
        for (int k = 1; k < nb; ++k) {
            const T ck = coef[k * 3 + ch];
            sx = sx + dYx[k] * ck;
            sy = sy + dYy[k] * ck;
            sz = sz + dYz[k] * ck;
        }
        dd[0] = dd[0] + sx * dl[ch];
        dd[1] = dd[1] + sy * dl[ch];
        dd[2] = dd[2] + sz * dl[ch];
    }
    T dm[3];
    dnormalize(raw_dir, dd, dm);
    gmean[0] += dm[0]; gmean[1] += dm[1]; gmean[2] += dm[2];
}

template <class T>
PHT_HD void sh_backward(int degree, const T raw_dir[3], const T d[3], const T* coef,
                        uint32_t clamp_bits, const T dl_rgb[3], T* gsh, T gmean[3]) {
    sh_backward_coef(degree, d, clamp_bits, dl_rgb, gsh);
    sh_backward_dir(degree, raw_dir, d, coef, clamp_bits, dl_rgb, gmean);
}

// ---------------------------------------------------------------- per-pixel blend term
// The blend evaluates each pair's Gaussian in base 2: with the conic scaled once per splat,
// (a2, b2, c2) = (-log2(e)/2 a, -log2(e) b, -log2(e)/2 c), the exponent
// p2 = a2 dx^2 + 2 (b2 / 2) dx dy + c2 dy^2 = log2(e) * power, and G = 2^p2 = exp(power).
// Same value as the reference's exp(power) (geometry.cuh:26-33, render_forward.cu:380-388);
// fewer operations per pair (one multiply per conic term is moved out of the pair loop, and the
// exponential needs no ln 2 reduction).
template <class T> struct Log2Consts {
    static constexpr T neg_half_log2e = T(-0.72134752044448170368);
    static constexpr T neg_log2e = T(-1.44269504088896340736);
};
template <class T> PHT_HD void conic_log2(T a, T b, T c, T& a2, T& b2, T& c2) {
    a2 = Log2Consts<T>::neg_half_log2e * a;
    b2 = Log2Consts<T>::neg_log2e * b;
    c2 = Log2Consts<T>::neg_half_log2e * c;
}
template <class T> PHT_HD T pair_power2(T a2, T b2, T c2, T dx, T dy) {
    return fma_(dx, fma_(a2, dx, b2 * dy), (c2 * dy) * dy);
}
// Skip bound in base-2 units: alpha = op 2^p2 >= 1/255 needs p2 >= log2(1/255) = -7.994 (op <= 1).
constexpr float kSkipPower2 = -8.0f;

// 2^x for x in [-8, 0] (or NaN): x = k + r, |r| <= 1/2, a degree-5 polynomial for 2^r (fitted
// to minimise the relative error on [-1/2, 1/2]: 9.2e-8 relative, <= 2.03 ulp in float Horner
// form), exact scaling by 2^k. NaN propagates. (A degree-6 fit, 0.94 ulp, costs one more fused
// multiply-add per evaluated pair in both blends.)
PHT_HD float pexp2f_poly(float r) {
    float p = 0.001326469355262816f;
    p = fma_(p, r, 0.00967151578515768f);
    p = fma_(p, r, 0.05550733953714371f);
    p = fma_(p, r, 0.24022242426872253f);
    p = fma_(p, r, 0.6931470036506653f);
    return fma_(p, r, 1.0f);
}
PHT_HD float pexp2f_blend(float x) {
    const float kf = rintf(x);
    const float p = pexp2f_poly(x - kf);
    const int k = x == x ? int(kf) : 0;
    return p * u2f(uint32_t(k + 127) << 23);
}
PHT_HD float exp2_blend_(float x) { return pexp2f_blend(x); }
PHT_HD double exp2_blend_(double x) { return exp2(x); }

// Alpha of one (instance, pixel) pair as in blend_tile (render_forward.cu:380-388), from the
// splat's base-2 conic. Returns false when the pair is skipped (exponent > 0 or alpha < 1/255);
// G = exp(power).
template <class T>
PHT_HD bool pair_alpha2(T a2, T b2, T c2, T op, T dx, T dy, T alpha_clip, T& alpha, T& G) {
    const T p2 = pair_power2(a2, b2, c2, dx, dy);
    if (p2 > T(0)) return false;
    if (p2 < T(kSkipPower2)) return false;
    G = exp2_blend_(p2);
    alpha = min_(alpha_clip, op * G);
    return !(alpha < K<T>::alpha_floor());
}
// The same from the splat's conic (a, b, c): the oracle's form.
template <class T>
PHT_HD bool pair_alpha(T ca, T cb, T cc, T op, T dx, T dy, T alpha_clip, T& alpha, T& G) {
    T a2, b2, c2;
    conic_log2(ca, cb, cc, a2, b2, c2);
    return pair_alpha2(a2, b2, c2, op, dx, dy, alpha_clip, alpha, G);
}

// ---------------------------------------------------------------- per-Gaussian backward
// Pinhole RGB chain of geo::splat_backward<PinholeRGB=true> (geometry.cuh:422-912 with no
// ray-plane / normal gradients: dL_dVrk = 0, dL_dr = 0, dL_du = dL_dv = dL_dz = 0),
// detail::cov3d_backward (:317-379) and add_mean2d_gradient (:385-417).
// Inputs: blend gradients d_conic (xyz conic, w opacity), d_mean2d. Outputs are added.
template <class T> struct SplatGrad {
    T mean[3];
    T scale[3];   // wrt activated (filtered) scale
    T rot[4];     // wrt normalised quaternion
    T opacity;    // wrt activated opacity
};

// The d(transform) -> d(mean3d) terms of one view and the mean2d -> mean3d term
// (geometry.cuh:826-903 pinhole, 385-399).
template <class T>
PHT_HD void splat_backward_pinhole_mean(const T mean[3], const pht_camera& cam, const float* view, T hx, T hy,
                                        const T t[3], T txc, T tyc, T x_mul, T y_mul, const M3<T>& Tm,
                                        const M3<T>& Vrk, const M3<T>& W, T dL_da, T dL_db, T dL_dc,
                                        const T d_mean2d[2], SplatGrad<T>& io) {
    (void)mean;
    (void)cam;
    const T dT00 = T(2) * ((Tm.at(0, 0) * Vrk.at(0, 0) + Tm.at(0, 1) * Vrk.at(0, 1)) + Tm.at(0, 2) * Vrk.at(0, 2)) * dL_da +
                   ((Tm.at(1, 0) * Vrk.at(0, 0) + Tm.at(1, 1) * Vrk.at(0, 1)) + Tm.at(1, 2) * Vrk.at(0, 2)) * dL_db;
    const T dT01 = T(2) * ((Tm.at(0, 0) * Vrk.at(1, 0) + Tm.at(0, 1) * Vrk.at(1, 1)) + Tm.at(0, 2) * Vrk.at(1, 2)) * dL_da +
                   ((Tm.at(1, 0) * Vrk.at(1, 0) + Tm.at(1, 1) * Vrk.at(1, 1)) + Tm.at(1, 2) * Vrk.at(1, 2)) * dL_db;
    const T dT02 = T(2) * ((Tm.at(0, 0) * Vrk.at(2, 0) + Tm.at(0, 1) * Vrk.at(2, 1)) + Tm.at(0, 2) * Vrk.at(2, 2)) * dL_da +
                   ((Tm.at(1, 0) * Vrk.at(2, 0) + Tm.at(1, 1) * Vrk.at(2, 1)) + Tm.at(1, 2) * Vrk.at(2, 2)) * dL_db;
    const T dT10 = T(2) * ((Tm.at(1, 0) * Vrk.at(0, 0) + Tm.at(1, 1) * Vrk.at(0, 1)) + Tm.at(1, 2) * Vrk.at(0, 2)) * dL_dc +
                   ((Tm.at(0, 0) * Vrk.at(0, 0) + Tm.at(0, 1) * Vrk.at(0, 1)) + Tm.at(0, 2) * Vrk.at(0, 2)) * dL_db;
    const T dT11 = T(2) * ((Tm.at(1, 0) * Vrk.at(1, 0) + Tm.at(1, 1) * Vrk.at(1, 1)) + Tm.at(1, 2) * Vrk.at(1, 2)) * dL_dc +
                   ((Tm.at(0, 0) * Vrk.at(1, 0) + Tm.at(0, 1) * Vrk.at(1, 1)) + Tm.at(0, 2) * Vrk.at(1, 2)) * dL_db;
    const T dT12 = T(2) * ((Tm.at(1, 0) * Vrk.at(2, 0) + Tm.at(1, 1) * Vrk.at(2, 1)) + Tm.at(1, 2) * Vrk.at(2, 2)) * dL_dc +
                   ((Tm.at(0, 0) * Vrk.at(2, 0) + Tm.at(0, 1) * Vrk.at(2, 1)) + Tm.at(0, 2) * Vrk.at(2, 2)) * dL_db;
    const T dJ00 = (W.at(0, 0) * dT00 + W.at(0, 1) * dT01) + W.at(0, 2) * dT02;
    const T dJ02 = (W.at(2, 0) * dT00 + W.at(2, 1) * dT01) + W.at(2, 2) * dT02;
    const T dJ11 = (W.at(1, 0) * dT10 + W.at(1, 1) * dT11) + W.at(1, 2) * dT12;
    const T dJ12 = (W.at(2, 0) * dT10 + W.at(2, 1) * dT11) + W.at(2, 2) * dT12;
    const T tz = T(1) / t[2];
    const T tz2 = tz * tz;
    const T tz3 = tz2 * tz;
    const T dt[3] = {
        x_mul * (-hx * tz2 * dJ02),
        y_mul * (-hy * tz2 * dJ12),
        ((-hx * tz2 * dJ00 - hy * tz2 * dJ11) + ((T(1) + x_mul) * hx * txc) * tz3 * dJ02) +
            ((T(1) + y_mul) * hy * tyc) * tz3 * dJ12};
    T gm[3];
    xform_dir_transpose(dt, view, gm);
    io.mean[0] += gm[0]; io.mean[1] += gm[1]; io.mean[2] += gm[2];

    // mean2d -> mean3d (geometry.cuh:385-399), unclamped point
    {
        const T rz = T(1) / (t[2] + T(1.0e-7));
        const T sx = t[0] * rz, sy = t[1] * rz;
        const T dmx = d_mean2d[0], dmy = d_mean2d[1];
        io.mean[0] += (hx * (T(view[0]) - sx * T(view[2])) * dmx + hy * (T(view[1]) - sy * T(view[2])) * dmy) * rz;
        io.mean[1] += (hx * (T(view[4]) - sx * T(view[6])) * dmx + hy * (T(view[5]) - sy * T(view[6])) * dmy) * rz;
        io.mean[2] += (hx * (T(view[8]) - sx * T(view[10])) * dmx + hy * (T(view[9]) - sy * T(view[10])) * dmy) * rz;
    }
}

// splat_backward_pinhole in two parts. The per-view part adds the view's d(2D conic) and
// d(mean2d) terms to io.mean / io.opacity and its d(3D covariance) (six values, dcov) to dcov_sum;
// splat_backward_cov_chain takes a d(3D covariance) to d(scale) and d(rotation). The chain is
// linear, so a Gaussian seen in several views can sum dcov over the views and chain once.
template <class T>
PHT_HD void splat_backward_pinhole_view(const T mean[3], const Activated<T>& act, const pht_camera& cam,
                                        const pht_settings& s, const T d_conic[4], const T d_mean2d[2],
                                        SplatGrad<T>& io, T dcov_sum[6]) {
    const float* view = cam.world_to_camera;
    const T hx = T(cam.fx), hy = T(cam.fy);
    T t[3];
    xform_point(mean, view, t);
    const T limx = T(1.3) * (T(cam.width) / (T(2) * hx));
    const T limy = T(1.3) * (T(cam.height) / (T(2) * hy));
    const T u0 = t[0] / t[2], v0 = t[1] / t[2];
    const T txc = min_(limx, max_(-limx, u0)) * t[2];
    const T tyc = min_(limy, max_(-limy, v0)) * t[2];
    const T x_mul = (u0 < -limx || u0 > limx) ? T(0) : T(1);
    const T y_mul = (v0 < -limy || v0 > limy) ? T(0) : T(1);
    const T tz2c = t[2] * t[2];
    M3<T> J;
    J.m[0][0] = hx / t[2]; J.m[0][1] = T(0); J.m[0][2] = T(0);
    J.m[1][0] = T(0); J.m[1][1] = hy / t[2]; J.m[1][2] = T(0);
    J.m[2][0] = -(hx * txc) / tz2c; J.m[2][1] = -(hy * tyc) / tz2c; J.m[2][2] = T(0);
    const M3<T> W = world_rotation_transposed<T>(view);
    const M3<T> Tm = mul(W, J);
    const T sm = T(s.scale_modifier);
    const T sl[3] = {sm * act.scale[0], sm * act.scale[1], sm * act.scale[2]};
    const M3<T> R = quat_rotation<T>(act.rot);
    M3<T> SR;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) SR.m[i][j] = sl[i] * R.m[i][j];
    const M3<T> M = mul(SR, Tm);
    const M3<T> cov2D = mul(transpose(M), M);
    const M3<T> Vrk = mul(transpose(SR), SR);
    unsigned min_id = 0;
    if (sl[1] < sl[min_id]) min_id = 1;
    if (sl[2] < sl[min_id]) min_id = 2;

    const T ks = T(s.kernel_size);
    T coef = T(1);
    if (s.compensate_2d) {
        const T det_0 = max_(T(1.0e-6), cov2D.m[0][0] * cov2D.m[1][1] - cov2D.m[0][1] * cov2D.m[0][1]);
        const T det_1 = max_(T(1.0e-6), (cov2D.m[0][0] + ks) * (cov2D.m[1][1] + ks) -
                                            cov2D.m[0][1] * cov2D.m[0][1]);
        coef = sqrt_(det_0 / det_1);
    }
    // conic -> 2D covariance (geometry.cuh:759-809)
    const T a = cov2D.m[0][0] + ks, b = cov2D.m[0][1], c = cov2D.m[1][1] + ks;
    const T denom = a * c - b * b;
    const T denom2inv = T(1) / (denom * denom + T(1.0e-7));
    T dL_da = T(0), dL_db = T(0), dL_dc = T(0);
    T dcov[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    const Cap<T> cp = cap_policy(a, b, c, cam.width, cam.height);
    if (denom2inv != T(0) || cp.q < T(1)) {
        if (s.fp64_conic && sizeof(T) == 4 && !(denom > T(kIllConditioned) * (a * c))) {
            conic_backward_f64(a, b, c, d_conic, dL_da, dL_db, dL_dc);
        } else {
            dL_da = denom2inv * ((-c * c * d_conic[0] + T(2) * b * c * d_conic[1]) + (denom - a * c) * d_conic[2]);
            dL_dc = denom2inv * ((-a * a * d_conic[2] + T(2) * a * b * d_conic[1]) + (denom - a * c) * d_conic[0]);
            dL_db = denom2inv * T(2) * ((b * c * d_conic[0] - (denom + T(2) * b * b) * d_conic[1]) + a * b * d_conic[2]);
        }
        if (cp.q < T(1)) {
            T g3[3];
            conic_backward_capped(a, b, c, cp, d_conic, g3);
            dL_da = g3[0]; dL_db = g3[1]; dL_dc = g3[2];
        }
        io.opacity += d_conic[3] * coef;
        const T t00 = Tm.at(0, 0), t10 = Tm.at(1, 0), t01 = Tm.at(0, 1), t11 = Tm.at(1, 1);
        const T t02 = Tm.at(0, 2), t12 = Tm.at(1, 2);
        dcov[0] = (t00 * t00 * dL_da + t00 * t10 * dL_db) + t10 * t10 * dL_dc;
        dcov[3] = (t01 * t01 * dL_da + t01 * t11 * dL_db) + t11 * t11 * dL_dc;
        dcov[5] = (t02 * t02 * dL_da + t02 * t12 * dL_db) + t12 * t12 * dL_dc;
        dcov[1] = (T(2) * t00 * t01 * dL_da + (t00 * t11 + t01 * t10) * dL_db) + T(2) * t10 * t11 * dL_dc;
        dcov[2] = (T(2) * t00 * t02 * dL_da + (t00 * t12 + t02 * t10) * dL_db) + T(2) * t10 * t12 * dL_dc;
        dcov[4] = (T(2) * t02 * t01 * dL_da + (t01 * t12 + t02 * t11) * dL_db) + T(2) * t11 * t12 * dL_dc;
    }

    for (int k = 0; k < 6; ++k) dcov_sum[k] = dcov_sum[k] + dcov[k];
    (void)min_id;

    // dL/dT -> dL/dJ -> dL/dt (geometry.cuh:826-903, pinhole)
    splat_backward_pinhole_mean(mean, cam, view, hx, hy, t, txc, tyc, x_mul, y_mul, Tm, Vrk, W, dL_da, dL_db, dL_dc,
                                d_mean2d, io);
}

// dSigma -> d(scale, rotation) (geometry.cuh:317-379), dL_dr = 0
template <class T>
PHT_HD void splat_backward_cov_chain(const Activated<T>& act, const pht_settings& s, const T dcov[6],
                                     SplatGrad<T>& io) {
    const T sm = T(s.scale_modifier);
    const T sl[3] = {sm * act.scale[0], sm * act.scale[1], sm * act.scale[2]};
    const M3<T> R = quat_rotation<T>(act.rot);
    {
        const T dsig[3][3] = {{dcov[0], T(0.5) * dcov[1], T(0.5) * dcov[2]},
                              {T(0.5) * dcov[1], dcov[3], T(0.5) * dcov[4]},
                              {T(0.5) * dcov[2], T(0.5) * dcov[4], dcov[5]}};
        M3<T> Ms; // S R with the local scales
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Ms.m[i][j] = sl[i] * R.m[i][j];
        M3<T> dM;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                T acc = T(0);
                for (int k = 0; k < 3; ++k) acc = acc + T(2) * Ms.m[i][k] * dsig[k][j];
                dM.m[i][j] = acc;
            }
        M3<T> dMt = transpose(dM);
        const T gs0 = (R.m[0][0] * dMt.m[0][0] + R.m[0][1] * dMt.m[1][0]) + R.m[0][2] * dMt.m[2][0];
        const T gs1 = (R.m[1][0] * dMt.m[0][1] + R.m[1][1] * dMt.m[1][1]) + R.m[1][2] * dMt.m[2][1];
        const T gs2 = (R.m[2][0] * dMt.m[0][2] + R.m[2][1] * dMt.m[1][2]) + R.m[2][2] * dMt.m[2][2];
        io.scale[0] += gs0 * sm;
        io.scale[1] += gs1 * sm;
        io.scale[2] += gs2 * sm;
        for (int r = 0; r < 3; ++r) {
            dMt.m[r][0] = dMt.m[r][0] * sl[0];
            dMt.m[r][1] = dMt.m[r][1] * sl[1];
            dMt.m[r][2] = dMt.m[r][2] * sl[2];
        }
        const T r_ = act.rot[0], x = act.rot[1], y = act.rot[2], z = act.rot[3];
        const T a01 = dMt.at(0, 1), a10 = dMt.at(1, 0);
        const T a02 = dMt.at(0, 2), a20 = dMt.at(2, 0);
        const T a12 = dMt.at(1, 2), a21 = dMt.at(2, 1);
        const T a00 = dMt.at(0, 0), a11 = dMt.at(1, 1), a22 = dMt.at(2, 2);
        io.rot[0] += (T(2) * z * (a01 - a10) + T(2) * y * (a20 - a02)) + T(2) * x * (a12 - a21);
        io.rot[1] += ((T(2) * y * (a10 + a01) + T(2) * z * (a20 + a02)) + T(2) * r_ * (a12 - a21)) -
                     T(4) * x * (a22 + a11);
        io.rot[2] += ((T(2) * x * (a10 + a01) + T(2) * r_ * (a20 - a02)) + T(2) * z * (a12 + a21)) -
                     T(4) * y * (a22 + a00);
        io.rot[3] += ((T(2) * r_ * (a01 - a10) + T(2) * x * (a20 + a02)) + T(2) * y * (a12 + a21)) -
                     T(4) * z * (a11 + a00);
    }
}

// One view, whole: the per-view part and the chain of its own d(3D covariance).
template <class T>
PHT_HD void splat_backward_pinhole(const T mean[3], const Activated<T>& act, const pht_camera& cam,
                                   const pht_settings& s, const T d_conic[4], const T d_mean2d[2],
                                   SplatGrad<T>& io) {
    T dcov[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    splat_backward_pinhole_view(mean, act, cam, s, d_conic, d_mean2d, io, dcov);
    splat_backward_cov_chain(act, s, dcov, io);
}

// Activation chain to raw parameters (update_structure_adam_row grad part,
// render_backward.cu:650-678 == chain_gradient_kernel, cuda_ops.cu:304-348), no regularisers.
template <class T>
PHT_HD void chain_raw(const Activated<T>& act, bool has_filter, T filter, const SplatGrad<T>& g,
                      T g_log_scale[3], T g_quat[4], T& g_logit) {
    const T f2 = has_filter ? filter * filter : T(0);
    for (int k = 0; k < 3; ++k) {
        const T fs = act.scale[k], rs = act.raw_scale[k];
        g_log_scale[k] = g.scale[k] * rs * rs / fs + g.opacity * act.opacity * f2 / (fs * fs);
    }
    const T qdot = ((act.rot[0] * g.rot[0] + act.rot[1] * g.rot[1]) + act.rot[2] * g.rot[2]) + act.rot[3] * g.rot[3];
    for (int k = 0; k < 4; ++k) g_quat[k] = act.inv_norm * (g.rot[k] - act.rot[k] * qdot);
    g_logit = g.opacity * act.opacity * (T(1) - act.sigmoid);
}

// ---------------------------------------------------------------- Adam element
// adam_element / adam_scalar (photara/src/splat/cuda_ops.cu:1291-1334,
// render_backward.cu:525-550). prefix_semantics selects the active-prefix SH variant
// (cuda_ops.cu:1506-1546): no clamping, non-finite moments leave the parameter untouched.
PHT_HD void adam_update(float& p, float& m, float& v, float g, float lr, float beta1, float beta2,
                        float c1, float c2, float eps, float lo, float hi, bool prefix_semantics) {
    const float prev = p;
    if (!finite_(prev) || !finite_(g)) {
        m = 0.0f;
        v = 0.0f;
        if (prefix_semantics) p = finite_(prev) ? prev : 0.0f;
        else p = finite_(prev) ? fminf(fmaxf(prev, lo), hi) : fminf(fmaxf(0.0f, lo), hi);
        return;
    }
    const float nm = beta1 * m + (1.0f - beta1) * g;
    const float nv = beta2 * v + ((1.0f - beta2) * g) * g;
    if (!finite_(nm) || !finite_(nv)) {
        m = 0.0f;
        v = 0.0f;
        if (!prefix_semantics) p = fminf(fmaxf(prev, lo), hi);
        return;
    }
    m = nm;
    v = nv;
    const float cand = prev - (lr * (nm / c1)) / (sqrtf(nv / c2) + eps);
    const float upd = finite_(cand) ? cand : prev;
    p = prefix_semantics ? upd : fminf(fmaxf(upd, lo), hi);
}

// log(1 + x): log(1 + x) directly for x >= 1; below, with u = fl(1 + x),
// log1p(x) = log(u) * x / (u - 1), exact for u == 1 (Goldberg, "What every computer scientist
// should know about floating-point arithmetic", 1991, thm 4).
PHT_HD float plog1pf(float x) {
    if (x != x || x == u2f(0x7f800000u)) return x;
    if (x >= 1.0f) return plogf(1.0f + x); // the rounding of 1 + x moves log by < 2^-24 / ln 2 ulp
    const float u = 1.0f + x;
    if (u == 1.0f) return x;
    return plogf(u) * (x / (u - 1.0f));
}

// exp(x) - 1: e^x - 1 for |x| >= 1/2; below, with u = fl(exp(x)), expm1(x) = (u - 1) * x / log(u)
// (Kahan's correction; the rounding of u cancels between numerator and denominator), x itself
// when u == 1.
PHT_HD float pexpm1f(float x) {
    if (x != x) return x;
    const float u = pexpf(x);
    if (!(fabsf(x) < 0.5f)) return u - 1.0f; // |e^x - 1| >= 0.39 |e^x| there: no cancellation
    if (u == 1.0f) return x;
    const float um1 = u - 1.0f;
    if (um1 == -1.0f) return -1.0f;
    if (!finite_(u)) return u;
    return um1 * (x / plogf(u));
}

// ---------------------------------------------------------------- quantised SH Adam state
// The reference trainer's SH optimizer state (sh_adam_quant_column / sh_adam_quant_warp_row,
// third_party/splat_drender/include/splat_drender/sh_adam_quant.cuh:25-234). Per coefficient it
// keeps the normalised Adam update u = (m / c1) / (sqrt(v / c2) + eps) in fp16 and an 8-bit code
// of log_s = log1p(sqrt(v) / 1e-15), linear between the row's bounds; DC columns [0, 3) and the
// rest [3, stride) each own a (min, max) pair. The raw moments are recovered from u with the
// previous step's corrections (c_{t-1} = (c_t - (1 - beta)) / beta). The codec constants and
// steps are a storage format, kept as the reference has them.
constexpr float kShqEps = 1e-15f;
constexpr float kShqInvEps = 1.0f / kShqEps;
constexpr float kShqInvQ = 1.0f / 255.0f;

PHT_HD float shq_decode(uint32_t code, float lo, float hi) { return lo + (hi - lo) * (float(code) * kShqInvQ); }
// Codes of a group are (log_s - lo) * scale rounded half away from zero and clamped to [0, 255],
// with scale = 255 / max(hi - lo, 1e-15) taken once per group.
PHT_HD float shq_scale(float lo, float hi) { return 255.0f / fmaxf(hi - lo, kShqEps); }
PHT_HD uint32_t shq_encode(float log_s, float lo, float scale) {
    return uint32_t(fminf(fmaxf(roundf((log_s - lo) * scale), 0.0f), 255.0f));
}
PHT_HD float shq_second(float log_s) {
    const float s = kShqEps * pexpm1f(log_s);
    return s * s;
}
PHT_HD float shq_log(float v) { return plog1pf(sqrtf(fmaxf(v, 0.0f)) * kShqInvEps); }

struct ShqStep {
    float lr, rest_lr, beta1, beta2, c1, c2, eps, regularization;
    int active;
};

// The step's per-coefficient constants, taken once: the previous step's corrections
// c_{t-1} = max((c_t - (1 - beta)) / beta, 0) and the reciprocals of the corrections.
struct ShqConst {
    ShqStep q;
    float pc1, inv_pc2, inv_c1, inv_c2;
    bool has_prev2, has_c;
};
PHT_HD ShqConst shq_const(const ShqStep& q) {
    ShqConst k;
    k.q = q;
    const float pc1 = q.beta1 > 0.0f ? fmaxf((q.c1 - (1.0f - q.beta1)) / q.beta1, 0.0f) : 0.0f;
    const float pc2 = q.beta2 > 0.0f ? fmaxf((q.c2 - (1.0f - q.beta2)) / q.beta2, 0.0f) : 0.0f;
    k.pc1 = pc1;
    k.has_prev2 = pc2 > 0.0f;
    k.inv_pc2 = k.has_prev2 ? 1.0f / pc2 : 0.0f;
    k.has_c = q.c1 > 0.0f && q.c2 > 0.0f;
    k.inv_c1 = k.has_c ? 1.0f / q.c1 : 0.0f;
    k.inv_c2 = k.has_c ? 1.0f / q.c2 : 0.0f;
    return k;
}
// (m / c1) / (sqrt(v / c2) + eps), 0 without valid corrections
PHT_HD float shq_normalised(const ShqConst& k, float m, float v) {
    return k.has_c ? (m * k.inv_c1) / (sqrtf(fmaxf(v, 0.0f) * k.inv_c2) + k.q.eps) : 0.0f;
}

// One coefficient: p (parameter) and u (stored first, decoded) are updated in place; returns the
// new log_s. log_s_old is the coefficient's decoded code under the row's old bounds.
PHT_HD float shq_column(const ShqConst& k, int col, float& p, float grad, float& u, float log_s_old) {
    const ShqStep& q = k.q;
    float log_s = log_s_old;
    float g2 = shq_second(log_s);
    float g1 = 0.0f;
    if (k.pc1 > 0.0f) g1 = (u * ((k.has_prev2 ? sqrtf(fmaxf(g2, 0.0f) * k.inv_pc2) : 0.0f) + q.eps)) * k.pc1;
    if (col >= q.active) { // inactive: keep the moments, renormalise u with this step's corrections
        if (!finite_(g1)) g1 = 0.0f;
        if (!finite_(log_s)) {
            log_s = 0.0f;
            g1 = 0.0f;
            g2 = 0.0f;
        }
        const float e = shq_normalised(k, g1, g2);
        u = finite_(e) ? e : 0.0f;
        return log_s;
    }
    const float prev = p;
    if (q.regularization > 0.0f && col >= 3 && finite_(prev)) grad += q.regularization * prev;
    if (!finite_(prev) || !finite_(grad)) {
        u = 0.0f;
        p = finite_(prev) ? prev : 0.0f;
        return 0.0f;
    }
    const float m = q.beta1 * g1 + (1.0f - q.beta1) * grad;
    const float v = q.beta2 * g2 + (1.0f - q.beta2) * grad * grad;
    if (!finite_(m) || !finite_(v)) {
        u = 0.0f;
        return 0.0f;
    }
    const float nrm = shq_normalised(k, m, v);
    const float cand = prev - (col >= 3 ? q.rest_lr : q.lr) * nrm;
    p = finite_(cand) ? cand : prev;
    u = nrm;
    const float ls = shq_log(v);
    return finite_(ls) ? ls : 0.0f;
}

// A row's stored bounds pair, with a non-finite pair read as (0, 0).
PHT_HD void shq_bounds(const float* b, int col, float& lo, float& hi) {
    const int o = col < 3 ? 0 : 2;
    lo = b[o];
    hi = b[o + 1];
    if (!finite_(lo) || !finite_(hi)) lo = hi = 0.0f;
}

// ---------------------------------------------------------------- reduced precision
PHT_HD float bf16_to_f32(uint16_t h) { return u2f(uint32_t(h) << 16); }
PHT_HD uint16_t f32_to_bf16(float f) { // round to nearest even, NaN kept quiet
    uint32_t u = f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
PHT_HD float fp16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1fu) return u2f(sign | 0x7f800000u | (m << 13));
    if (e == 0) {
        if (m == 0) return u2f(sign);
        const float f = float(m) * 5.9604644775390625e-08f; // 2^-24
        return sign ? -f : f;
    }
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}
PHT_HD uint16_t f32_to_fp16(float f) { // round to nearest even
    const uint32_t u = f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t au = u & 0x7fffffffu;
    if (au > 0x7f800000u) return uint16_t(sign | 0x7e00u);
    if (au >= 0x477ff000u) return uint16_t(sign | 0x7c00u); // >= 65520 rounds to inf
    if (au < 0x38800000u) { // subnormal half (or zero)
        const float af = u2f(au) * 16777216.0f; // scale by 2^24: value in half-ulp units
        const float r = rintf(af);
        return uint16_t(sign | uint32_t(r));
    }
    uint32_t e = (au >> 23) - 112u;
    uint32_t m = au & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return uint16_t(sign | h);
}
PHT_HD float lp_to_f32(uint16_t h, int fmt) { return fmt == PHT_LP_FP16 ? fp16_to_f32(h) : bf16_to_f32(h); }
PHT_HD uint16_t f32_to_lp(float f, int fmt) { return fmt == PHT_LP_FP16 ? f32_to_fp16(f) : f32_to_bf16(f); }

// ---------------------------------------------------------------- densification (per row)
// Shared by the device kernels and the CPU oracle, so both compute the same bits.

PHT_HD float psigmoidf(float x) { return 1.0f / (1.0f + pexpf(-x)); }
// log2 through plogf (~2 ulp), and x^y for x >= 0 through exp(y log x).
PHT_HD float plog2f(float x) { return plogf(x) * 1.44269502162933350f; }
PHT_HD float ppowf(float x, float y) { return x == 0.0f ? 0.0f : pexpf(y * plogf(x)); }
// logit(p) = log(p / (1 - p)).
PHT_HD float plogitf(float p) { return plogf(p / (1.0f - p)); }

// accumulate_densification_kernel (photara/src/splat/cuda_ops.cu:1548-1606). One row's statistics
// held in registers while the observations of several views are folded in, in order.
struct DensifyRow {
    float gradient, count, max_screen_radius, priority, max_geometry_gradient, view_support, accumulated_error;
    int32_t first_view;
};
PHT_HD DensifyRow densify_load_row(const pht_densify_stats& s, int64_t i) {
    return DensifyRow{s.gradient[i], s.count[i], s.max_screen_radius[i], s.priority[i], s.max_geometry_gradient[i],
                      s.view_support[i], s.accumulated_error[i], s.first_view[i]};
}
PHT_HD void densify_store_row(const pht_densify_stats& s, int64_t i, const DensifyRow& r) {
    s.gradient[i] = r.gradient;
    s.count[i] = r.count;
    s.max_screen_radius[i] = r.max_screen_radius;
    s.priority[i] = r.priority;
    s.max_geometry_gradient[i] = r.max_geometry_gradient;
    s.view_support[i] = r.view_support;
    s.accumulated_error[i] = r.accumulated_error;
    s.first_view[i] = r.first_view;
}
// One observation of one row: radius, visibility and refine weight of the view, the optional
// geometry gradient / image error / normalised size, under the observation's rules.
PHT_HD void densify_observe(DensifyRow& s, const pht_densify_observation& o, int32_t view_index,
                            float inverse_resolution, int32_t radius, float visibility, float rw, bool has_geom,
                            float geom, bool has_err, float err, bool has_size, float size) {
    if (radius <= 0 || (o.require_contribution_visibility && visibility <= 0.0f)) return;
    const float raw = finite_(rw) ? fmaxf(rw, 0.0f) : 0.0f;
    if (has_geom && finite_(geom)) s.max_geometry_gradient = fmaxf(s.max_geometry_gradient, geom);
    const float w = o.step_score_power == 1.0f ? raw : ppowf(raw, o.step_score_power);
    s.gradient = o.use_maximum ? fmaxf(s.gradient, w) : s.gradient + w;
    s.count += 1.0f;
    if (has_err && finite_(err)) s.accumulated_error += fmaxf(err, 0.0f);
    if (view_index >= 0) {
        if (s.first_view == 0) {
            s.first_view = view_index + 1;
            s.view_support = 1.0f;
        } else if (s.first_view != view_index + 1) {
            s.view_support = 2.0f;
        }
    }
    const float screen = has_size ? size : float(radius) * inverse_resolution;
    s.max_screen_radius = fmaxf(s.max_screen_radius, screen);
    const float thr = o.oversize_screen_threshold;
    s.priority += thr > 0.0f ? fmaxf(plog2f(screen / thr), 0.0f) : w * (1.0f + screen);
}
PHT_HD void densify_accumulate_row(DensifyRow& s, const pht_densify_observation& o, float inverse_resolution,
                                   int64_t i) {
    const int32_t radius = o.radii[i];
    if (radius <= 0) return;
    const bool has_geom = o.geometry_gradient != nullptr, has_err = o.image_error != nullptr;
    const bool has_size = o.normalized_size != nullptr;
    densify_observe(s, o, o.view_index, inverse_resolution, radius, o.visibility[i], o.refine_weight[i], has_geom,
                    has_geom ? o.geometry_gradient[i] : 0.0f, has_err, has_err ? o.image_error[i] : 0.0f, has_size,
                    has_size ? o.normalized_size[i] : 0.0f);
}

// adc_plus_prune_kernel (photara/src/splat/cuda_ops.cu:263-302): row i. Returns opacity.
PHT_HD float densify_prune_row(const float* means, const float* log_scales, const float* quats, const float* logits,
                               const float* sh, int sh_stride, int64_t i, float minimum_opacity,
                               float maximum_bounds, const float center[3], bool& keep, bool& hard) {
    const float logit = logits[i];
    const float opacity = psigmoidf(logit);
    bool bad = !finite_(logit);
    float max_scale = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float m = means[3 * i + k], ls = log_scales[3 * i + k];
        bad = bad || !finite_(m) || !finite_(ls);
        max_scale = fmaxf(max_scale, pexpf(ls));
    }
    for (int k = 0; k < 4; ++k) bad = bad || !finite_(quats[4 * i + k]);
    if (sh != nullptr) // (a caller may check the SH values itself and pass nullptr)
        for (int k = 0; k < sh_stride; ++k) bad = bad || !finite_(sh[i * sh_stride + k]);
    const bool outside = fabsf(means[3 * i] - center[0]) > maximum_bounds ||
                         fabsf(means[3 * i + 1] - center[1]) > maximum_bounds ||
                         fabsf(means[3 * i + 2] - center[2]) > maximum_bounds;
    hard = bad || outside || max_scale > maximum_bounds;
    keep = !hard && opacity >= minimum_opacity;
    return opacity;
}

// Selection keys: larger key = selected first; key 0 = not eligible.
// Total order of floats as unsigned integers (negative values flipped, positive values above).
PHT_HD uint32_t float_order_bits(float v) {
    const uint32_t u = f2u(v);
    return u ^ ((u & 0x80000000u) ? 0xffffffffu : 0x80000000u);
}
PHT_HD float float_from_order_bits(uint32_t o) {
    return u2f((o & 0x80000000u) ? (o ^ 0x80000000u) : ~o);
}
// Largest / smallest value first, ties to the lower row (a stable sort's front).
PHT_HD uint64_t select_value_key(float v, bool largest, uint32_t i) {
    const uint32_t o = float_order_bits(v);
    return (uint64_t(largest ? o : ~o) << 32) | uint64_t(0xffffffffu - i);
}
// splitmix64's output function (Steele, Lea, Flood 2014): the generator's i-th output for seed s
// is mix(s + 0x9e3779b97f4a7c15 * i).
PHT_HD uint64_t splitmix64_mix(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
// Gumbel top-k key of igs_weighted_sample_cuda (photara/src/splat/igs_selection_cuda.cu:31-70):
// score = log(w) - log(-log(u)) for positive finite w (subnormal w through its integer bits),
// u uniform in (0, 1) from splitmix64; ties to the higher row.
PHT_HD uint64_t gumbel_key(float w, uint64_t seed, uint32_t i) {
    const uint32_t bits = f2u(w);
    if (bits == 0u || bits >= 0x7f800000u) return 0;
    const uint64_t r = splitmix64_mix(seed + 0x9e3779b97f4a7c15ull * (uint64_t(i) + 1));
    const float u = fminf(1.0f - 1e-7f, fmaxf(1e-7f, (float(uint32_t(r >> 41)) + 0.5f) * 1.1920928955078125e-07f));
    const float log_w = bits < 0x00800000u ? plogf(float(bits)) - 149.0f * 0.6931471805599453f : plogf(w);
    const float score = log_w - plogf(-plogf(u));
    return (uint64_t(float_order_bits(score)) << 32) | uint64_t(i);
}

// Quaternion rotation v' = v + w t + q x t, t = 2 q x v, of the normalised scalar-first q.
PHT_HD void quat_rotate(const float* raw, const float v[3], float out[3]) {
    const float inv = 1.0f / sqrtf(fmaxf(((raw[0] * raw[0] + raw[1] * raw[1]) + raw[2] * raw[2]) + raw[3] * raw[3],
                                         1e-20f));
    const float w = raw[0] * inv, x = raw[1] * inv, y = raw[2] * inv, z = raw[3] * inv;
    const float tx = 2.0f * (y * v[2] - z * v[1]);
    const float ty = 2.0f * (z * v[0] - x * v[2]);
    const float tz = 2.0f * (x * v[1] - y * v[0]);
    out[0] = (v[0] + w * tx) + (y * tz - z * ty);
    out[1] = (v[1] + w * ty) + (z * tx - x * tz);
    out[2] = (v[2] + w * tz) + (x * ty - y * tx);
}

// split_gaussians_kernel (photara/src/splat/cuda_ops.cu:1720-1845): the parent row and the child
// row both hold the parent's parameters on entry; both receive the split parameters.
PHT_HD void densify_split_row(float* means, float* log_scales, float* logits, const float* quats, int64_t parent,
                              int64_t child, const float sample[3], float screen, bool has_screen, int mode,
                              float minimum_opacity, float split_at_screen_size, float split_opacity_k) {
    const float* ls = log_scales + 3 * parent;
    float local[3] = {0.0f, 0.0f, 0.0f}, delta[3] = {0.0f, 0.0f, 0.0f};
    if (mode == PHT_SPLIT_DENSE_TANGENT) {
        // tangent-plane split: offsets along local x / y, both shrink by 1/sqrt(2), normal kept
        local[0] = (pexpf(ls[0]) * sample[0]) * 0.5f;
        local[1] = (pexpf(ls[1]) * sample[1]) * 0.5f;
        delta[0] = delta[1] = plogf(0.70710678118654752f);
    } else if (mode == PHT_SPLIT_ADC_COVARIANCE) {
        // rank-one split along the major axis that keeps the mixture covariance
        float scale2[3], max2 = 0.0f;
        int major = 0;
        for (int k = 0; k < 3; ++k) {
            const float s = pexpf(ls[k]);
            scale2[k] = s * s;
            if (scale2[k] > max2) { max2 = scale2[k]; major = k; }
        }
        const float standard_k = 0.70710678118654752f;
        const float sc = has_screen ? fmaxf(screen, 1e-6f) : 1e-6f;
        const float max_k = split_at_screen_size > 0.0f ? fminf(standard_k, split_at_screen_size / sc) : standard_k;
        for (int k = 0; k < 3; ++k) {
            const float kk = k == major ? max_k : 1.0f;
            local[k] = sqrtf(fmaxf(1.0f - kk * kk, 0.0f)) * sqrtf(scale2[k]);
            delta[k] = plogf(fmaxf(kk, 1e-12f));
        }
    } else if (mode == PHT_SPLIT_IGS_RANDOM) {
        // one shared variate scales every axis; only the largest axis halves
        int largest = 0;
        if (ls[1] > ls[largest]) largest = 1;
        if (ls[2] > ls[largest]) largest = 2;
        for (int k = 0; k < 3; ++k) {
            local[k] = pexpf(ls[k]) * sample[0];
            delta[k] = k == largest ? plogf(0.5f) : 0.0f;
        }
    } else { // PHT_SPLIT_LONG_AXIS: offset of half the major scale along it; major halves, others x0.85
        int largest = 0;
        for (int k = 1; k < 3; ++k)
            if (ls[k] > ls[largest]) largest = k;
        const float offset = 0.5f * pexpf(ls[largest]);
        for (int k = 0; k < 3; ++k) {
            local[k] = k == largest ? offset : 0.0f;
            delta[k] = k == largest ? plogf(0.5f) : plogf(0.85f);
        }
    }
    float off[3];
    quat_rotate(quats + 4 * parent, local, off);
    for (int k = 0; k < 3; ++k) {
        const float c = means[3 * parent + k];
        means[3 * parent + k] = c - off[k];
        means[3 * child + k] = c + off[k];
        log_scales[3 * parent + k] += delta[k];
        log_scales[3 * child + k] += delta[k];
    }
    float logit;
    if (mode == PHT_SPLIT_LONG_AXIS) {
        // a' = k / (1 + exp(-l) - k): two children approximately composite back to the parent
        const float den = 1.0f + pexpf(-logits[parent]) - split_opacity_k;
        const float mapped = den > 1e-6f ? fminf(fmaxf(split_opacity_k / den, 1e-6f), 1.0f - 1e-6f) : 1.0f - 1e-6f;
        logit = plogitf(mapped);
    } else {
        // 1 - (1 - a)^p with p = 1/2 (two children reproduce the parent's alpha exactly) or
        // 1/sqrt(2) (covariance mode), clamped to [min, 1 - min]
        const float opacity = psigmoidf(logits[parent]);
        const float base = fmaxf(1.0f - opacity, 0.0f);
        const float powered = mode == PHT_SPLIT_ADC_COVARIANCE ? ppowf(base, 0.70710678118654752f) : sqrtf(base);
        const float revised = fminf(fmaxf(1.0f - powered, minimum_opacity), 1.0f - minimum_opacity);
        logit = plogitf(revised);
    }
    logits[parent] = logit;
    logits[child] = logit;
}

// adc_decay_kernel (photara/src/splat/cuda_ops.cu:1847-1861): row i.
PHT_HD void densify_decay_row(float* log_scales, float* logits, int64_t i, float opacity_decay, float log_scale_decay) {
    const float op = fminf(fmaxf(psigmoidf(logits[i]) - opacity_decay, 1e-12f), 1.0f - 1e-12f);
    logits[i] = plogitf(op);
    for (int k = 0; k < 3; ++k) log_scales[3 * i + k] += log_scale_decay;
}

// IGS parent scores of retained row r (densification_igs.cpp:439-497).
struct IgsRow {
    float priority;     // priority / max(count, 1)
    bool candidate;     // count > 0
    bool positive;      // candidate && priority > 0
    bool growth;        // candidate && gradient > gradient threshold
    bool oversized;     // candidate && max screen radius > screen threshold
};
PHT_HD IgsRow igs_row(const pht_densify_stats& s, int64_t src, float gradient_threshold, float screen_threshold) {
    IgsRow r;
    const float c = s.count[src];
    r.priority = s.priority[src] / fmaxf(c, 1.0f);
    r.candidate = c > 0.0f;
    r.positive = r.candidate && r.priority > 0.0f;
    r.growth = r.candidate && s.gradient[src] > gradient_threshold;
    r.oversized = r.candidate && s.max_screen_radius[src] > screen_threshold;
    return r;
}
PHT_HD void igs_weights_row(const pht_densify_stats& s, int64_t src, float opacity, float median, bool allow_growth,
                            float gradient_threshold, float screen_threshold, float& replacement, float& growth,
                            float& oversize) {
    const IgsRow r = igs_row(s, src, gradient_threshold, screen_threshold);
    const float edge = fminf((r.priority > 0.0f ? r.priority : 0.0f) / median, 10.0f) * 0.25f + 1.0f;
    replacement = r.candidate ? opacity * edge : 0.0f;
    growth = allow_growth && r.growth ? (s.gradient[src] * edge) * (r.oversized ? 2.0f : 1.0f) : 0.0f;
    oversize = allow_growth && r.oversized ? s.max_screen_radius[src] * edge : 0.0f;
}

} // namespace pht

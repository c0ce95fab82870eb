// CPU oracle of the 3DGS training step — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline load this library, and
// only as the checker / the timed CPU baseline; the product path never links or calls it.
//
// It restates the reference rasterizer's algorithm on the CPU, templated on the arithmetic
// type: float is the parity oracle (same per-element formulas as the kernels, so keys,
// ranges, contributor counts and images must match bit-exactly), double is the accuracy
// oracle that FP32/BF16 results are measured against.
//   - per (gaussian, view) preprocess: render_forward.cu:30-139 (activation, EWA, conic,
//     radius, AccuTile count, depth key, SH colour)
//   - instance list: emit_packed_instances + stable radix sort (render_forward.cu:184-194,
//     pipeline.cu:240-249): order (tile, depth bits, gaussian index), done with std::stable_sort
//   - ranges: extract_packed_ranges (render_forward.cu:196-212)
//   - per-pixel front-to-back blend: blend_tile (render_forward.cu:332-416, 508-525)
//   - per-pixel back-to-front backward from the final transmittance: the classic 3DGS
//     backward, the same closed-form terms as blend_bucket_backward (render_backward.cu:383-489)
//   - per-Gaussian backward + activation chain + Adam: render_backward.cu:650-862
//   - L1 loss gradient: cuda_ops.cu:886-898
// Functions whose results none of the pins in tests/ cover are listed as "parity unpinned"
// in the test suite; every function here is pinned by a closed form, a brute-force or an
// independent numpy restatement (tests/test_oracle_*.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../photara_studio_amd/csrc/splat_math.h"

using namespace pht;

namespace {

template <class T> struct SplatRec {
    int32_t g, v;
    T mean2d[2];
    T conic[3];
    T opacity;
    T rgb[3];
    uint32_t clamp_bits;
    uint32_t depth_key;
    uint32_t tiles;
    bool fp64;
};

struct Views {
    std::vector<ViewInfo> info;
    int64_t tiles = 0;
    int64_t pixels = 0;
};

Views make_views(const pht_camera* cams, int n, int ts) {
    Views vs;
    for (int i = 0; i < n; ++i) {
        ViewInfo vi{};
        vi.cam = cams[i];
        vi.grid_x = (cams[i].width + ts - 1) / ts;
        vi.grid_y = (cams[i].height + ts - 1) / ts;
        vi.tile_base = int32_t(vs.tiles);
        vi.tile_count = vi.grid_x * vi.grid_y;
        vi.pixel_base = vs.pixels;
        vs.tiles += vi.tile_count;
        vs.pixels += int64_t(cams[i].width) * cams[i].height;
        vs.info.push_back(vi);
    }
    return vs;
}

template <class T> T load_in(const float* f, const uint16_t* lp, int fmt, int64_t i) {
    if (fmt != PHT_LP_NONE && lp) return T(lp_to_f32(lp[i], fmt));
    return T(f[i]);
}

// Optional double-precision parameters for the accuracy oracle (finite differences need inputs
// that are not rounded to float). When set, the double instantiation reads these arrays.
struct Inputs64 {
    const double* means = nullptr;
    const double* log_scales = nullptr;
    const double* quaternions = nullptr;
    const double* opacity_logits = nullptr;
    const double* sh = nullptr;
};
Inputs64 g_in64;

template <class T> T mean_in(const pht_gaussians& g, int64_t j) {
    if (sizeof(T) == 8 && g_in64.means) return T(g_in64.means[j]);
    return T(g.means[j]);
}

template <class T> Activated<T> activate_row(const pht_gaussians& g, int64_t i) {
    T ls[3], q[4];
    T logit;
    if (sizeof(T) == 8 && g_in64.log_scales) {
        for (int k = 0; k < 3; ++k) ls[k] = T(g_in64.log_scales[3 * i + k]);
        for (int k = 0; k < 4; ++k) q[k] = T(g_in64.quaternions[4 * i + k]);
        logit = T(g_in64.opacity_logits[i]);
    } else {
        for (int k = 0; k < 3; ++k) ls[k] = load_in<T>(g.log_scales, g.log_scales_lp, g.lp_format, 3 * i + k);
        for (int k = 0; k < 4; ++k) q[k] = T(g.quaternions[4 * i + k]);
        logit = load_in<T>(g.opacity_logits, g.opacity_logits_lp, g.lp_format, i);
    }
    const bool hf = g.filter_3d != nullptr;
    return activate<T>(ls, q, logit, hf, hf ? T(g.filter_3d[i]) : T(0));
}

template <class T> void load_sh(const pht_gaussians& g, const pht_settings& s, int64_t i, T* coef) {
    const int nb = sh_count(s.active_sh_degree);
    for (int k = 0; k < nb; ++k)
        for (int ch = 0; ch < 3; ++ch)
            coef[k * 3 + ch] = (sizeof(T) == 8 && g_in64.sh)
                                   ? T(g_in64.sh[(i * s.sh_bases + k) * 3 + ch])
                                   : load_in<T>(g.sh, g.sh_lp, g.lp_format, (i * s.sh_bases + k) * 3 + ch);
}

// Per (gaussian, view) preprocess. Returns true when visible (touches >= 1 tile).
template <class T>
bool preprocess_one(const pht_gaussians& g, const pht_settings& s, const ViewInfo& vi, int64_t i,
                    SplatRec<T>& rec, int32_t& radius_out) {
    radius_out = 0;
    const T mean[3] = {mean_in<T>(g, 3 * i), mean_in<T>(g, 3 * i + 1), mean_in<T>(g, 3 * i + 2)};
    const Activated<T> act = activate_row<T>(g, i);
    Projected<T> p;
    if (!project_pinhole(mean, act, vi.cam, s, p)) return false;
    const uint32_t touched = enumerate_tiles(p.mean2d[0], p.mean2d[1], p.conic, p.opacity, vi.grid_x,
                                             vi.grid_y, s.tile_size, [](int, int) {});
    if (touched == 0) return false;
    rec.mean2d[0] = p.mean2d[0]; rec.mean2d[1] = p.mean2d[1];
    for (int k = 0; k < 3; ++k) rec.conic[k] = p.conic[k];
    rec.opacity = p.opacity;
    rec.tiles = touched;
    rec.depth_key = f2u(float(p.depth));
    rec.fp64 = p.fp64;
    rec.clamp_bits = 0;
    if (g.colors_precomp) {
        for (int ch = 0; ch < 3; ++ch) rec.rgb[ch] = T(g.colors_precomp[3 * i + ch]);
    } else {
        T coef[48], d[3], raw[3];
        load_sh<T>(g, s, i, coef);
        view_dir(mean, vi.cam.position, d, raw);
        sh_eval(s.active_sh_degree, d, coef, rec.rgb, rec.clamp_bits);
    }
    radius_out = p.radius;
    return true;
}

template <class T> struct State {
    Views views;
    std::vector<SplatRec<T>> splats;     // visible splats in (view, gaussian) order
    std::vector<int64_t> splat_of;       // [view*N + g] -> splat index or -1
    std::vector<uint64_t> keys;          // sorted
    std::vector<uint32_t> vals;          // sorted: splat index
    std::vector<uint32_t> ranges;        // [2*tiles]
    std::vector<T> final_T;              // [pixels]
    std::vector<uint32_t> n_contrib;     // [pixels]
    std::vector<T> color;                // per view [3,H,W], concatenated
    std::vector<T> alpha;                // per view [H,W], concatenated
    int64_t fp64_count = 0;
};

template <class T>
void run_forward(const pht_gaussians& g, const pht_camera* cams, int n_views, const pht_settings& s,
                 State<T>& st, pht_view_outputs* outs, int threads) {
    const int64_t N = g.count;
    st.views = make_views(cams, n_views, s.tile_size);
    const Views& vs = st.views;
    st.splat_of.assign(size_t(N) * n_views, -1);
    std::vector<SplatRec<T>> dense(size_t(N) * n_views);
    std::vector<uint8_t> vis(size_t(N) * n_views, 0);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < N; ++i) {
        for (int v = 0; v < n_views; ++v) {
            SplatRec<T>& r = dense[size_t(v) * N + i];
            int32_t radius = 0;
            const bool ok = preprocess_one<T>(g, s, vs.info[v], i, r, radius);
            if (outs && outs[v].radii) outs[v].radii[i] = ok ? radius : 0;
            if (ok) {
                r.g = int32_t(i);
                r.v = v;
                vis[size_t(v) * N + i] = 1;
            }
        }
    }
    st.splats.clear();
    st.fp64_count = 0;
    for (int v = 0; v < n_views; ++v)
        for (int64_t i = 0; i < N; ++i)
            if (vis[size_t(v) * N + i]) {
                st.splat_of[size_t(v) * N + i] = int64_t(st.splats.size());
                st.splats.push_back(dense[size_t(v) * N + i]);
                st.fp64_count += dense[size_t(v) * N + i].fp64 ? 1 : 0;
            }
    // Instance list: (global_tile << 32 | depth) keys emitted in (view, gaussian) order, then a
    // stable sort by key == the reference's order (tile, depth bits, gaussian index).
    std::vector<std::pair<uint64_t, uint32_t>> inst;
    for (size_t si = 0; si < st.splats.size(); ++si) {
        const SplatRec<T>& r = st.splats[si];
        const ViewInfo& vi = vs.info[r.v];
        enumerate_tiles(r.mean2d[0], r.mean2d[1], r.conic, r.opacity, vi.grid_x, vi.grid_y, s.tile_size,
                        [&](int tx, int ty) {
                            const uint64_t tile = uint64_t(vi.tile_base) + uint64_t(ty) * vi.grid_x + tx;
                            inst.push_back({(tile << 32) | r.depth_key, uint32_t(si)});
                        });
    }
    std::stable_sort(inst.begin(), inst.end(),
                     [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    st.keys.resize(inst.size());
    st.vals.resize(inst.size());
    for (size_t k = 0; k < inst.size(); ++k) {
        st.keys[k] = inst[k].first;
        st.vals[k] = inst[k].second;
    }
    st.ranges.assign(size_t(2 * vs.tiles), 0);
    for (size_t k = 0; k < st.keys.size(); ++k) {
        const uint64_t tile = st.keys[k] >> 32;
        if (k == 0 || (st.keys[k - 1] >> 32) != tile) st.ranges[2 * tile] = uint32_t(k);
        st.ranges[2 * tile + 1] = uint32_t(k + 1);
    }
    // Blend, one tile at a time (tiles are independent).
    st.final_T.assign(size_t(vs.pixels), T(1));
    st.n_contrib.assign(size_t(vs.pixels), 0);
    st.color.assign(size_t(3 * vs.pixels), T(0));
    st.alpha.assign(size_t(vs.pixels), T(0));
    const T floorT = K<T>::transmittance_floor();
    const T clip = T(s.alpha_clip);
    const int ts = s.tile_size;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int64_t tile = 0; tile < vs.tiles; ++tile) {
        int v = 0;
        while (v + 1 < n_views && vs.info[v + 1].tile_base <= tile) ++v;
        const ViewInfo& vi = vs.info[v];
        const int local = int(tile - vi.tile_base);
        const int tx = local % vi.grid_x, ty = local / vi.grid_x;
        const uint32_t b = st.ranges[2 * tile], e = st.ranges[2 * tile + 1];
        const int W = vi.cam.width, H = vi.cam.height;
        const int64_t P = int64_t(W) * H;
        for (int py = ty * ts; py < std::min(H, (ty + 1) * ts); ++py)
            for (int px = tx * ts; px < std::min(W, (tx + 1) * ts); ++px) {
                T Tr = T(1), C[3] = {T(0), T(0), T(0)};
                uint32_t last = 0;
                for (uint32_t k = b; k < e; ++k) {
                    const SplatRec<T>& r = st.splats[st.vals[k]];
                    const T dx = r.mean2d[0] - T(px), dy = r.mean2d[1] - T(py);
                    T alpha, G;
                    if (!pair_alpha(r.conic[0], r.conic[1], r.conic[2], r.opacity, dx, dy, clip, alpha, G))
                        continue;
                    const T test_t = Tr * (T(1) - alpha);
                    if (test_t < floorT) break;
                    const T a_t = alpha * Tr;
                    for (int ch = 0; ch < 3; ++ch) C[ch] = fma_(r.rgb[ch], a_t, C[ch]);
                    Tr = test_t;
                    last = k - b + 1;
                    if (outs && outs[v].visibility) outs[v].visibility[r.g] = 1.0f;
                }
                const int64_t pid = int64_t(py) * W + px;
                st.final_T[size_t(vi.pixel_base + pid)] = Tr;
                st.n_contrib[size_t(vi.pixel_base + pid)] = last;
                for (int ch = 0; ch < 3; ++ch)
                    st.color[size_t(3 * vi.pixel_base + ch * P + pid)] = C[ch] + Tr * T(vi.cam.background[ch]);
                st.alpha[size_t(vi.pixel_base + pid)] = T(1) - Tr;
                if (outs) {
                    for (int ch = 0; ch < 3; ++ch)
                        outs[v].color[ch * P + pid] = float(C[ch] + Tr * T(vi.cam.background[ch]));
                    outs[v].alpha[pid] = float(T(1) - Tr);
                }
            }
    }
}

// Accuracy oracle on the float run's discrete decisions ("matched FP64"): the tile lists, the
// per-pair skip test (alpha < 1/255, power outside [-5.6, 0]), the early termination and the SH
// clamp flags are those of the float run `f`; every value (projection, colour, alpha, the
// composited sums, the gradients) is computed in double. The image is piecewise smooth in the
// inputs, with jumps where a pair crosses the 1/255 cut-off; comparing on the same piece isolates
// the arithmetic error from such threshold flips.
// With `double_decisions` (precision 66, "order-matched FP64") only the float run's tile lists and
// blend order are kept; which pairs pass the 1/255 cut-off and where a pixel terminates are decided
// in double. Against the plain FP64 oracle this isolates the order of splats whose depths are
// within float rounding of each other (the reference sorts by float32 depth bits).
void run_forward_matched(const pht_gaussians& g, const pht_settings& s, const State<float>& f,
                         State<double>& st, int threads, bool double_decisions = false) {
    st.views = f.views;
    st.splat_of = f.splat_of;
    st.keys = f.keys;
    st.vals = f.vals;
    st.ranges = f.ranges;
    st.n_contrib = f.n_contrib;
    st.fp64_count = f.fp64_count;
    st.splats.assign(f.splats.size(), SplatRec<double>{});
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t si = 0; si < int64_t(f.splats.size()); ++si) {
        const SplatRec<float>& r32 = f.splats[size_t(si)];
        SplatRec<double> r{};
        int32_t radius = 0;
        if (!preprocess_one<double>(g, s, st.views.info[size_t(r32.v)], r32.g, r, radius)) {
            // culled in double only (a near-plane or zero-tile flip): carry the float values
            r.mean2d[0] = r32.mean2d[0]; r.mean2d[1] = r32.mean2d[1];
            for (int k = 0; k < 3; ++k) { r.conic[k] = r32.conic[k]; r.rgb[k] = r32.rgb[k]; }
            r.opacity = r32.opacity;
        }
        r.g = r32.g;
        r.v = r32.v;
        r.clamp_bits = r32.clamp_bits;
        r.depth_key = r32.depth_key;
        r.tiles = r32.tiles;
        r.fp64 = r32.fp64;
        st.splats[size_t(si)] = r;
    }
    const Views& vs = st.views;
    const int n_views = int(vs.info.size());
    st.final_T.assign(size_t(vs.pixels), 1.0);
    st.color.assign(size_t(3 * vs.pixels), 0.0);
    st.alpha.assign(size_t(vs.pixels), 0.0);
    const float clip32 = s.alpha_clip;
    const double clip = double(s.alpha_clip);
    const int ts = s.tile_size;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int64_t tile = 0; tile < vs.tiles; ++tile) {
        int v = 0;
        while (v + 1 < n_views && vs.info[v + 1].tile_base <= tile) ++v;
        const ViewInfo& vi = vs.info[v];
        const int local = int(tile - vi.tile_base);
        const int tx = local % vi.grid_x, ty = local / vi.grid_x;
        const uint32_t b = st.ranges[2 * tile], e = st.ranges[2 * tile + 1];
        const int W = vi.cam.width, H = vi.cam.height;
        const int64_t P = int64_t(W) * H;
        for (int py = ty * ts; py < std::min(H, (ty + 1) * ts); ++py)
            for (int px = tx * ts; px < std::min(W, (tx + 1) * ts); ++px) {
                float T32 = 1.0f;
                double Tr = 1.0, C[3] = {0.0, 0.0, 0.0};
                for (uint32_t k = b; k < e; ++k) {
                    if (double_decisions) {
                        const SplatRec<double>& r = st.splats[st.vals[k]];
                        double alpha, G;
                        if (!pair_alpha(r.conic[0], r.conic[1], r.conic[2], r.opacity, r.mean2d[0] - double(px),
                                        r.mean2d[1] - double(py), clip, alpha, G))
                            continue;
                        const double test = Tr * (1.0 - alpha);
                        if (test < K<double>::transmittance_floor()) break;
                        for (int ch = 0; ch < 3; ++ch) C[ch] = fma(r.rgb[ch], alpha * Tr, C[ch]);
                        Tr = test;
                        continue;
                    }
                    const SplatRec<float>& q = f.splats[f.vals[k]];
                    float a32, G32;
                    if (!pair_alpha(q.conic[0], q.conic[1], q.conic[2], q.opacity, q.mean2d[0] - float(px),
                                    q.mean2d[1] - float(py), clip32, a32, G32))
                        continue;
                    const float test32 = T32 * (1.0f - a32);
                    if (test32 < K<float>::transmittance_floor()) break;
                    T32 = test32;
                    const SplatRec<double>& r = st.splats[st.vals[k]];
                    double a2, b2, c2;
                    conic_log2(r.conic[0], r.conic[1], r.conic[2], a2, b2, c2);
                    const double alpha = std::min(
                        clip, r.opacity * exp2(pair_power2(a2, b2, c2, r.mean2d[0] - double(px), r.mean2d[1] - double(py))));
                    const double a_t = alpha * Tr;
                    for (int ch = 0; ch < 3; ++ch) C[ch] = fma(r.rgb[ch], a_t, C[ch]);
                    Tr = Tr * (1.0 - alpha);
                }
                const int64_t pid = int64_t(py) * W + px;
                st.final_T[size_t(vi.pixel_base + pid)] = Tr;
                for (int ch = 0; ch < 3; ++ch)
                    st.color[size_t(3 * vi.pixel_base + ch * P + pid)] = C[ch] + Tr * double(vi.cam.background[ch]);
                st.alpha[size_t(vi.pixel_base + pid)] = 1.0 - Tr;
            }
    }
}

// Per-splat blend gradients (the GradState of render_backward.cu): d_mean2d xy, d_conic xyz,
// d_opacity, d_color, refine.
template <class T> struct SplatGradAcc {
    T dmean[2], dconic[3], dop, dcol[3], refine;
};

// With `dec` set (matched FP64), which pairs contribute is decided by the float run `dec`.
template <class T>
void run_blend_backward(const pht_gaussians& g, const pht_settings& s, const State<T>& st,
                        const pht_view_grads* dl, std::vector<SplatGradAcc<T>>& acc, int threads,
                        const State<float>* dec = nullptr) {
    (void)g;
    const Views& vs = st.views;
    const int n_views = int(vs.info.size());
    // Per-instance contributions, reduced afterwards in instance order (deterministic).
    std::vector<SplatGradAcc<T>> per_inst(st.vals.size());
    std::memset((void*)per_inst.data(), 0, per_inst.size() * sizeof(SplatGradAcc<T>));
    const T clip = T(s.alpha_clip);
    const int ts = s.tile_size;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int64_t tile = 0; tile < vs.tiles; ++tile) {
        int v = 0;
        while (v + 1 < n_views && vs.info[v + 1].tile_base <= tile) ++v;
        const ViewInfo& vi = vs.info[v];
        const int local = int(tile - vi.tile_base);
        const int tx = local % vi.grid_x, ty = local / vi.grid_x;
        const uint32_t b = st.ranges[2 * tile];
        const int W = vi.cam.width, H = vi.cam.height;
        const int64_t P = int64_t(W) * H;
        const T bg[3] = {T(vi.cam.background[0]), T(vi.cam.background[1]), T(vi.cam.background[2])};
        for (int py = ty * ts; py < std::min(H, (ty + 1) * ts); ++py)
            for (int px = tx * ts; px < std::min(W, (tx + 1) * ts); ++px) {
                const int64_t pid = int64_t(py) * W + px;
                const T Tfin = st.final_T[size_t(vi.pixel_base + pid)];
                const uint32_t last = st.n_contrib[size_t(vi.pixel_base + pid)];
                if (last == 0) continue;
                const T gp[3] = {T(dl[v].grad_color[pid]), T(dl[v].grad_color[P + pid]),
                                 T(dl[v].grad_color[2 * P + pid])};
                const T ga = dl[v].grad_alpha ? T(dl[v].grad_alpha[pid]) : T(0);
                const T dL_dTfin = ((bg[0] * gp[0] + bg[1] * gp[1]) + bg[2] * gp[2]) - ga;
                const T refine_den = T(1) / max_(T(1) - Tfin, T(1.0e-5));
                T Tr = Tfin;
                T Ca[3] = {T(0), T(0), T(0)};
                for (int64_t k = int64_t(b) + last - 1; k >= int64_t(b); --k) {
                    const SplatRec<T>& r = st.splats[st.vals[size_t(k)]];
                    const T dx = r.mean2d[0] - T(px), dy = r.mean2d[1] - T(py);
                    T alpha, G;
                    if (dec) {
                        const SplatRec<float>& q = dec->splats[dec->vals[size_t(k)]];
                        float a32, G32;
                        if (!pair_alpha(q.conic[0], q.conic[1], q.conic[2], q.opacity, q.mean2d[0] - float(px),
                                        q.mean2d[1] - float(py), float(s.alpha_clip), a32, G32))
                            continue;
                        T a2, b2, c2;
                        conic_log2(r.conic[0], r.conic[1], r.conic[2], a2, b2, c2);
                        G = exp2_blend_(pair_power2(a2, b2, c2, dx, dy));
                        alpha = min_(clip, r.opacity * G);
                    } else if (!pair_alpha(r.conic[0], r.conic[1], r.conic[2], r.opacity, dx, dy, clip, alpha, G)) {
                        continue;
                    }
                    const T one_m = T(1) - alpha;
                    const T inv_1ma = T(1) / one_m;
                    Tr = Tr * inv_1ma;
                    const T w = alpha * Tr;
                    const T c_dot = (r.rgb[0] * gp[0] + r.rgb[1] * gp[1]) + r.rgb[2] * gp[2];
                    const T a_dot = ((Ca[0] * gp[0] + Ca[1] * gp[1]) + Ca[2] * gp[2]) * inv_1ma;
                    const T dL_dalpha = (Tr * c_dot - a_dot) - Tfin * inv_1ma * dL_dTfin;
                    const T dL_dG = r.opacity * dL_dalpha;
                    const T gdx = G * dx, gdy = G * dy;
                    const T dG_dx = -gdx * r.conic[0] - gdy * r.conic[1];
                    const T dG_dy = -gdy * r.conic[2] - gdx * r.conic[1];
                    const T ddx = dL_dG * dG_dx, ddy = dL_dG * dG_dy;
                    SplatGradAcc<T>& a = per_inst[size_t(k)];
                    a.dmean[0] += ddx;
                    a.dmean[1] += ddy;
                    a.dconic[0] += T(-0.5) * gdx * dx * dL_dG;
                    a.dconic[1] += T(-0.5) * gdx * dy * dL_dG;
                    a.dconic[2] += T(-0.5) * gdy * dy * dL_dG;
                    a.dop += G * dL_dalpha;
                    for (int ch = 0; ch < 3; ++ch) a.dcol[ch] += w * gp[ch];
                    const T rx = ddx * T(W), ry = ddy * T(H);
                    a.refine += sqrt_(rx * rx + ry * ry) * refine_den;
                    for (int ch = 0; ch < 3; ++ch) Ca[ch] = fma_(r.rgb[ch], w, Ca[ch]);
                }
            }
    }
    acc.assign(st.splats.size(), SplatGradAcc<T>{});
    std::memset((void*)acc.data(), 0, acc.size() * sizeof(SplatGradAcc<T>));
    for (size_t k = 0; k < st.vals.size(); ++k) {
        SplatGradAcc<T>& d = acc[st.vals[k]];
        const SplatGradAcc<T>& a = per_inst[k];
        d.dmean[0] += a.dmean[0]; d.dmean[1] += a.dmean[1];
        for (int c = 0; c < 3; ++c) { d.dconic[c] += a.dconic[c]; d.dcol[c] += a.dcol[c]; }
        d.dop += a.dop;
        d.refine += a.refine;
    }
}

template <class T> struct RowGrad {
    T mean[3], log_scale[3], quat[4], logit, sh[48], colors[3], refine;
};

// Per-Gaussian backward summed over views (gaussian_backward, render_backward.cu:708-832).
template <class T>
void gaussian_row_grad(const pht_gaussians& g, const pht_settings& s, const State<T>& st,
                       const std::vector<SplatGradAcc<T>>& acc, int64_t i, RowGrad<T>& out) {
    std::memset((void*)&out, 0, sizeof(out));
    const int64_t N = g.count;
    const int n_views = int(st.views.info.size());
    const T mean[3] = {mean_in<T>(g, 3 * i), mean_in<T>(g, 3 * i + 1), mean_in<T>(g, 3 * i + 2)};
    const Activated<T> act = activate_row<T>(g, i);
    SplatGrad<T> sg;
    std::memset((void*)&sg, 0, sizeof(sg));
    bool any = false;
    T coef[48];
    if (!g.colors_precomp) load_sh<T>(g, s, i, coef);
    for (int v = 0; v < n_views; ++v) {
        const int64_t si = st.splat_of[size_t(v) * N + i];
        if (si < 0) continue;
        any = true;
        const SplatGradAcc<T>& a = acc[size_t(si)];
        const ViewInfo& vi = st.views.info[v];
        const T dconic[4] = {a.dconic[0], a.dconic[1], a.dconic[2], a.dop};
        splat_backward_pinhole(mean, act, vi.cam, s, dconic, a.dmean, sg);
        if (g.colors_precomp) {
            for (int ch = 0; ch < 3; ++ch) out.colors[ch] += a.dcol[ch];
        } else {
            T d[3], raw[3];
            view_dir(mean, vi.cam.position, d, raw);
            sh_backward(s.active_sh_degree, raw, d, coef, st.splats[size_t(si)].clamp_bits, a.dcol,
                        out.sh, sg.mean);
        }
        out.refine += a.refine;
    }
    (void)any;
    for (int k = 0; k < 3; ++k) out.mean[k] = sg.mean[k];
    const bool hf = g.filter_3d != nullptr;
    chain_raw(act, hf, hf ? T(g.filter_3d[i]) : T(0), sg, out.log_scale, out.quat, out.logit);
}

template <class T, class O>
void write_grads(const pht_gaussians& g, const pht_settings& s, int64_t i, const RowGrad<T>& r,
                 O* means, O* log_scales, O* quats, O* logits, O* sh, O* colors, O* refine) {
    for (int k = 0; k < 3; ++k) {
        if (means) means[3 * i + k] = O(r.mean[k]);
        if (log_scales) log_scales[3 * i + k] = O(r.log_scale[k]);
        if (colors) colors[3 * i + k] = O(r.colors[k]);
    }
    for (int k = 0; k < 4; ++k)
        if (quats) quats[4 * i + k] = O(r.quat[k]);
    if (logits) logits[i] = O(r.logit);
    if (refine) refine[i] = O(r.refine);
    if (sh) {
        const int nb = sh_count(s.active_sh_degree);
        for (int k = 0; k < s.sh_bases; ++k)
            for (int ch = 0; ch < 3; ++ch)
                sh[(i * s.sh_bases + k) * 3 + ch] = k < nb ? O(r.sh[k * 3 + ch]) : O(0);
    }
}

// One row of the quantised SH Adam (sh_adam_quant_warp_row,
// third_party/splat_drender/include/splat_drender/sh_adam_quant.cuh:146-234), walked column by
// column as the reference's own CPU check does (photara/tests/splat_test.cpp:1334-1434).
inline void shq_row(const ShqStep& q, int stride, float* p, const float* grad, uint16_t* first, uint8_t* packed,
                    float* bounds) {
    float ls[64];
    const ShqConst kc = shq_const(q);
    for (int k = 0; k < stride; ++k) {
        float lo, hi;
        shq_bounds(bounds, k, lo, hi);
        float u = fp16_to_f32(first[k]);
        float pv = p[k];
        const bool act = k < q.active;
        ls[k] = shq_column(kc, k, pv, act ? grad[k] : 0.0f, u, shq_decode(packed[k], lo, hi));
        if (act) p[k] = pv;
        first[k] = f32_to_fp16(u);
    }
    float b[4] = {1e30f, -1e30f, 1e30f, -1e30f};
    for (int k = 0; k < stride; ++k) {
        const int o = k < 3 ? 0 : 2;
        b[o] = fminf(b[o], ls[k]);
        b[o + 1] = fmaxf(b[o + 1], ls[k]);
    }
    if (stride <= 3) b[2] = b[3] = 0.0f;
    for (int o = 0; o < 4; ++o) bounds[o] = b[o];
    const float sc[2] = {shq_scale(b[0], b[1]), shq_scale(b[2], b[3])};
    for (int k = 0; k < stride; ++k) {
        const int g = k < 3 ? 0 : 1;
        packed[k] = uint8_t(shq_encode(ls[k], b[2 * g], sc[g]));
    }
}

// One fused Adam row over all parameters (structure + active-prefix SH).
inline void adam_row(const pht_adam& a, const pht_settings& s, int64_t i, const float gm[3],
                     const float gls[3], const float gq[4], float glog, const float* gsh) {
    const float inf = u2f(0x7f800000u);
    for (int k = 0; k < 3; ++k) {
        const int64_t j = 3 * i + k;
        adam_update(a.means[j], a.m_means[j], a.v_means[j], gm[k], a.lr_means, a.beta1, a.beta2,
                    a.correction1, a.correction2, a.epsilon, -inf, inf, false);
        adam_update(a.log_scales[j], a.m_log_scales[j], a.v_log_scales[j], gls[k], a.lr_scales, a.beta1,
                    a.beta2, a.correction1, a.correction2, a.epsilon, a.min_log_scale, a.max_log_scale, false);
    }
    for (int k = 0; k < 4; ++k) {
        const int64_t j = 4 * i + k;
        adam_update(a.quaternions[j], a.m_quaternions[j], a.v_quaternions[j], gq[k], a.lr_quaternions,
                    a.beta1, a.beta2, a.correction1, a.correction2, a.epsilon, -inf, inf, false);
    }
    adam_update(a.opacity_logits[i], a.m_opacity_logits[i], a.v_opacity_logits[i], glog, a.lr_opacity,
                a.beta1, a.beta2, a.correction1, a.correction2, a.epsilon, -12.0f, 12.0f, false);
    if (a.sh && a.sh_first) {
        const int active = sh_count(s.active_sh_degree) * 3;
        const int stride = s.sh_bases * 3;
        const int64_t o = i * int64_t(stride);
        shq_row(ShqStep{a.lr_sh0, a.lr_sh_rest, a.beta1, a.beta2, a.correction1, a.correction2, a.epsilon,
                        a.sh_regularization, active},
                stride, a.sh + o, gsh, a.sh_first + o, a.sh_packed + o, a.sh_bounds + 4 * i);
    } else if (a.sh) {
        const int active = sh_count(s.active_sh_degree) * 3;
        for (int c = 0; c < active; ++c) {
            const int64_t j = i * int64_t(s.sh_bases) * 3 + c;
            adam_update(a.sh[j], a.m_sh[j], a.v_sh[j], gsh[c], c < 3 ? a.lr_sh0 : a.lr_sh_rest, a.beta1,
                        a.beta2, a.correction1, a.correction2, a.epsilon, -inf, inf, true);
        }
    }
    if (a.lp_format != PHT_LP_NONE) {
        for (int k = 0; k < 3; ++k)
            if (a.log_scales_lp) a.log_scales_lp[3 * i + k] = f32_to_lp(a.log_scales[3 * i + k], a.lp_format);
        if (a.opacity_logits_lp) a.opacity_logits_lp[i] = f32_to_lp(a.opacity_logits[i], a.lp_format);
        if (a.sh_lp)
            for (int c = 0; c < s.sh_bases * 3; ++c)
                a.sh_lp[i * int64_t(s.sh_bases) * 3 + c] = f32_to_lp(a.sh[i * int64_t(s.sh_bases) * 3 + c], a.lp_format);
    }
}

State<float> g_state32;
State<double> g_state64;
std::vector<SplatGradAcc<float>> g_acc32;
std::vector<SplatGradAcc<double>> g_acc64;
float g_alpha_clip = 0.99f; // alpha clip of the last forward (for the tile-free check)

template <class T>
void export_splat_grads(const State<T>& st, const std::vector<SplatGradAcc<T>>& acc, int64_t n, double* out) {
    const int nv = int(st.views.info.size());
    for (int64_t g = 0; g < n; ++g)
        for (int v = 0; v < nv; ++v) {
            double* o = out + (g * nv + v) * 10;
            const int64_t si = st.splat_of[size_t(v) * n + g];
            if (si < 0) {
                for (int k = 0; k < 10; ++k) o[k] = 0.0;
                continue;
            }
            const SplatGradAcc<T>& a = acc[size_t(si)];
            o[0] = a.dmean[0]; o[1] = a.dmean[1];
            o[2] = a.dconic[0]; o[3] = a.dconic[1]; o[4] = a.dconic[2];
            o[5] = a.dop; o[6] = a.dcol[0]; o[7] = a.dcol[1]; o[8] = a.dcol[2]; o[9] = a.refine;
        }
}

} // namespace

// Fused L1 + SSIM loss and gradient of one view (fused_ssim.cu:120-470, the fused-ssim
// formulation): separable 11-tap Gaussian moments with zero padding, SSIM partials on the valid
// window [5, W-5) x [5, H-5), the separable backward window, plus the L1 term. Float: the same
// operations in the same order as the GPU kernel; double: the accuracy reference. Returns the
// loss (summed in double); grad is overwritten.
const float kSsimWindow[11] = {0.001028380123898387f, 0.0075987582094967365f, 0.036000773310661316f,
                               0.10936068743467331f,  0.21300552785396576f,   0.26601171493530273f,
                               0.21300552785396576f,  0.10936068743467331f,   0.036000773310661316f,
                               0.0075987582094967365f, 0.001028380123898387f};

template <class T>
double l1_ssim(const float* color, const float* target, const float* mask, int W, int H, float ssim_w_f,
               float photometric_f, T* grad) {
    const int64_t P = int64_t(W) * H;
    const T ssim_w = T(ssim_w_f);
    // the normalisation is a float computed on the host (fused_ssim.cu:535-536), in both precisions
    const T norm = T(photometric_f / float(3 * (W - 10) * (H - 10)));
    const T chain_ssim = -ssim_w * norm;
    const T C1 = T(0.01f * 0.01f), C2 = T(0.03f * 0.03f);
    T w[11];
    for (int d = 0; d < 11; ++d) w[d] = T(kSsimWindow[d]);
    auto inside = [&](int x, int y) { return x >= 0 && y >= 0 && x < W && y < H; };
    auto valid = [&](int x, int y) { return x >= 5 && x < W - 5 && y >= 5 && y < H - 5; };
    const size_t np = static_cast<size_t>(P);
    std::vector<T> X(np), Y(np), hm(5 * np), dmap(3 * np), hd(3 * np);
    double loss = 0.0;
    for (int ch = 0; ch < 3; ++ch) {
        for (int64_t p = 0; p < P; ++p) {
            const float m = mask ? mask[p] : 1.0f;
            X[size_t(p)] = T(mask ? color[ch * P + p] * m : color[ch * P + p]);
            Y[size_t(p)] = T(mask ? target[ch * P + p] * m : target[ch * P + p]);
        }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0), s4 = T(0);
                for (int d = 0; d < 11; ++d) {
                    const int xx = x + d - 5;
                    const T a = inside(xx, y) ? X[size_t(y) * W + xx] : T(0);
                    const T b = inside(xx, y) ? Y[size_t(y) * W + xx] : T(0);
                    s0 += a * w[d];
                    s1 += a * a * w[d];
                    s2 += b * w[d];
                    s3 += b * b * w[d];
                    s4 += a * b * w[d];
                }
                const size_t i = size_t(y) * W + x;
                hm[i] = s0; hm[P + i] = s1; hm[2 * P + i] = s2; hm[3 * P + i] = s3; hm[4 * P + i] = s4;
            }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                T m[5] = {T(0), T(0), T(0), T(0), T(0)};
                for (int d = 0; d < 11; ++d) {
                    const int yy = y + d - 5;
                    for (int k = 0; k < 5; ++k)
                        m[k] += (yy >= 0 && yy < H ? hm[size_t(k) * P + size_t(yy) * W + x] : T(0)) * w[d];
                }
                const size_t i = size_t(y) * W + x;
                T d0 = T(0), d1 = T(0), d2 = T(0);
                if (valid(x, y)) {
                    const T mu1 = m[0], mu2 = m[2];
                    const T mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2;
                    const T sigma1_sq = m[1] - mu1_sq, sigma2_sq = m[3] - mu2_sq, sigma12 = m[4] - mu1 * mu2;
                    const T A = mu1_sq + mu2_sq + C1;
                    const T B = sigma1_sq + sigma2_sq + C2;
                    const T Cc = T(2) * mu1 * mu2 + C1;
                    const T D = T(2) * sigma12 + C2;
                    const T ab = A * B;
                    d0 = ((T(2) * mu2 * D) / ab - (T(2) * mu2 * Cc) / ab - (T(2) * mu1 * Cc * D) / (A * A * B) +
                          (T(2) * mu1 * Cc * D) / (A * B * B)) * chain_ssim;
                    d1 = (-Cc * D / (A * B * B)) * chain_ssim;
                    d2 = (T(2) * Cc / ab) * chain_ssim;
                    const T ssim = Cc * D / ab;
                    const T l1 = std::fabs(X[i] - Y[i]);
                    loss += double(norm * (ssim_w * (T(1) - ssim) + (T(1) - ssim_w) * l1));
                }
                dmap[i] = d0; dmap[P + i] = d1; dmap[2 * P + i] = d2;
            }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                T a[3] = {T(0), T(0), T(0)};
                for (int d = 0; d < 11; ++d) {
                    const int xx = x + d - 5;
                    for (int k = 0; k < 3; ++k)
                        a[k] += (inside(xx, y) ? dmap[size_t(k) * P + size_t(y) * W + xx] : T(0)) * w[d];
                }
                const size_t i = size_t(y) * W + x;
                hd[i] = a[0]; hd[P + i] = a[1]; hd[2 * P + i] = a[2];
            }
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                T g[3] = {T(0), T(0), T(0)};
                for (int d = 0; d < 11; ++d) {
                    const int yy = y + d - 5;
                    for (int k = 0; k < 3; ++k)
                        g[k] += (yy >= 0 && yy < H ? hd[size_t(k) * P + size_t(yy) * W + x] : T(0)) * w[d];
                }
                const size_t i = size_t(y) * W + x;
                const T xv = X[i], yv = Y[i];
                const T chain = valid(x, y) ? norm : T(0);
                const T sign = xv == yv ? T(0) : std::copysign(T(1), xv - yv);
                const T m = mask ? T(mask[i]) : T(1);
                grad[size_t(ch) * P + i] = (g[0] + T(2) * xv * g[1] + yv * g[2] + (T(1) - ssim_w) * sign * chain) * m;
            }
    }
    return loss;
}

extern "C" {

int oracle_abi_version(void) { return PHT_ABI_VERSION; }

// Double-precision parameters read by the precision-64 runs (NULL means: use the float inputs).
void oracle_set_inputs64(const double* means, const double* log_scales, const double* quaternions,
                         const double* opacity_logits, const double* sh) {
    g_in64.means = means;
    g_in64.log_scales = log_scales;
    g_in64.quaternions = quaternions;
    g_in64.opacity_logits = opacity_logits;
    g_in64.sh = sh;
}

// Deterministic math primitives, exported for the error-bound tests.
void oracle_pexpf(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = pexpf(x[i]);
}
void oracle_pexp2f_blend(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = pexp2f_blend(x[i]);
}
void oracle_plogf(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = plogf(x[i]);
}
void oracle_plog1pf(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = plog1pf(x[i]);
}
void oracle_pexpm1f(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = pexpm1f(x[i]);
}
// The blend's base-2 exponent of each (conic, offset) pair (conic_log2 + pair_power2).
void oracle_pair_power2(const float* conic3, const float* d2, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        float a2, b2, c2;
        conic_log2(conic3[3 * i], conic3[3 * i + 1], conic3[3 * i + 2], a2, b2, c2);
        out[i] = pair_power2(a2, b2, c2, d2[2 * i], d2[2 * i + 1]);
    }
}
void oracle_lp_roundtrip(const float* x, uint16_t* bits, float* back, int64_t n, int fmt) {
    for (int64_t i = 0; i < n; ++i) {
        bits[i] = f32_to_lp(x[i], fmt);
        back[i] = lp_to_f32(bits[i], fmt);
    }
}

// Per-(gaussian) projection of one view: mean2d(2), conic(3), opacity, depth, radius, tiles,
// rgb(3), clamp bits, fp64 flag — for the independent numpy restatement tests.
int oracle_project(int precision, const pht_gaussians* g, const pht_camera* cam, const pht_settings* s,
                   double* out16) {
    const ViewInfo vi = make_views(cam, 1, s->tile_size).info[0];
    for (int64_t i = 0; i < g->count; ++i) {
        double* o = out16 + 16 * i;
        for (int k = 0; k < 16; ++k) o[k] = 0.0;
        int32_t radius = 0;
        bool ok;
        if (precision == 64) {
            SplatRec<double> r{};
            ok = preprocess_one<double>(*g, *s, vi, i, r, radius);
            if (ok) {
                o[0] = r.mean2d[0]; o[1] = r.mean2d[1];
                o[2] = r.conic[0]; o[3] = r.conic[1]; o[4] = r.conic[2];
                o[5] = r.opacity; o[6] = double(u2f(r.depth_key)); o[8] = r.tiles;
                o[9] = r.rgb[0]; o[10] = r.rgb[1]; o[11] = r.rgb[2]; o[12] = r.clamp_bits; o[13] = r.fp64;
            }
        } else {
            SplatRec<float> r{};
            ok = preprocess_one<float>(*g, *s, vi, i, r, radius);
            if (ok) {
                o[0] = r.mean2d[0]; o[1] = r.mean2d[1];
                o[2] = r.conic[0]; o[3] = r.conic[1]; o[4] = r.conic[2];
                o[5] = r.opacity; o[6] = double(u2f(r.depth_key)); o[8] = r.tiles;
                o[9] = r.rgb[0]; o[10] = r.rgb[1]; o[11] = r.rgb[2]; o[12] = r.clamp_bits; o[13] = r.fp64;
            }
        }
        o[7] = radius;
        o[14] = ok ? 1.0 : 0.0;
    }
    return 0;
}

// Tiles of one projected splat in walk order (tile ids), for the brute-force coverage tests.
int oracle_enumerate(float mx, float my, const float* conic3, float op, int grid_x, int grid_y, int ts,
                     int32_t* tiles_out, int cap) {
    const float c[3] = {conic3[0], conic3[1], conic3[2]};
    int n = 0;
    enumerate_tiles(mx, my, c, op, grid_x, grid_y, ts, [&](int tx, int ty) {
        if (n < cap) tiles_out[n] = ty * grid_x + tx;
        ++n;
    });
    return n;
}

// Forward. precision 32: parity oracle (float outputs); 64: accuracy oracle (double results
// rounded to float in `outs`); 65: accuracy oracle on the float run's decisions (see
// run_forward_matched; `outs` is not written); 66: on the float run's tile lists and order only
// (images only: no backward). Keys/values/ranges/n_contrib/final_T of the float run are copied
// when the pointers are non-null.
int oracle_forward(int precision, const pht_gaussians* g, const pht_camera* cams, int n_views,
                   const pht_settings* s, pht_view_outputs* outs, uint64_t* keys, uint32_t* values,
                   int64_t key_cap, uint32_t* ranges, uint32_t* n_contrib, float* final_T,
                   pht_counts* counts, int threads) {
    g_alpha_clip = s->alpha_clip;
    if (precision == 65 || precision == 66) {
        run_forward<float>(*g, cams, n_views, *s, g_state32, nullptr, threads);
        run_forward_matched(*g, *s, g_state32, g_state64, threads, precision == 66);
        if (counts) {
            counts->visible = int64_t(g_state64.splats.size());
            counts->instances = int64_t(g_state64.keys.size());
            counts->fp64_conics = g_state64.fp64_count;
            counts->overflow = 0;
        }
        return 0;
    }
    if (precision == 64) {
        run_forward<double>(*g, cams, n_views, *s, g_state64, outs, threads);
        if (counts) {
            counts->visible = int64_t(g_state64.splats.size());
            counts->instances = int64_t(g_state64.keys.size());
            counts->fp64_conics = g_state64.fp64_count;
            counts->overflow = 0;
        }
        return 0;
    }
    State<float>& st = g_state32;
    run_forward<float>(*g, cams, n_views, *s, st, outs, threads);
    if (counts) {
        counts->visible = int64_t(st.splats.size());
        counts->instances = int64_t(st.keys.size());
        counts->fp64_conics = st.fp64_count;
        counts->overflow = 0;
    }
    const int64_t n = std::min<int64_t>(key_cap, int64_t(st.keys.size()));
    for (int64_t k = 0; k < n; ++k) {
        if (keys) keys[k] = st.keys[size_t(k)];
        if (values) values[k] = uint32_t(st.splats[st.vals[size_t(k)]].g);
    }
    if (ranges) std::copy(st.ranges.begin(), st.ranges.end(), ranges);
    if (n_contrib) std::copy(st.n_contrib.begin(), st.n_contrib.end(), n_contrib);
    if (final_T) std::copy(st.final_T.begin(), st.final_T.end(), final_T);
    return 0;
}

// SH colour of unit directions (n x 3) for one coefficient row (coef[k*3+ch]), in double; rgb
// before the max(., 0) clamp is returned in out[n x 3].
void oracle_sh_eval(int degree, const double* dirs, const double* coef, double* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        double Y[16];
        sh_basis(degree, d, Y);
        for (int ch = 0; ch < 3; ++ch) {
            double acc = Y[0] * coef[ch];
            for (int k = 1; k < sh_count(degree); ++k) acc = fma(Y[k], coef[k * 3 + ch], acc);
            out[3 * i + ch] = acc + 0.5;
        }
    }
}

// Tile-free render of the last float forward: every pixel walks all visible splats of its view in
// (depth bits, gaussian index) order with the same per-pair arithmetic. Equal to the tiled render
// bit for bit iff tile enumeration, sorting and ranges lose nothing.
int oracle_bruteforce(float* color_concat, float* alpha_concat, float* final_T) {
    const State<float>& st = g_state32;
    for (size_t vi_idx = 0; vi_idx < st.views.info.size(); ++vi_idx) {
        const ViewInfo& vi = st.views.info[vi_idx];
        std::vector<uint32_t> order;
        for (size_t si = 0; si < st.splats.size(); ++si)
            if (st.splats[si].v == int(vi_idx)) order.push_back(uint32_t(si));
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const SplatRec<float>& x = st.splats[a];
            const SplatRec<float>& y = st.splats[b];
            if (x.depth_key != y.depth_key) return x.depth_key < y.depth_key;
            return x.g < y.g;
        });
        const int W = vi.cam.width, H = vi.cam.height;
        const int64_t P = int64_t(W) * H;
#pragma omp parallel for schedule(dynamic, 8)
        for (int py = 0; py < H; ++py)
            for (int px = 0; px < W; ++px) {
                float Tr = 1.0f, C[3] = {0.0f, 0.0f, 0.0f};
                for (uint32_t si : order) {
                    const SplatRec<float>& r = st.splats[si];
                    const float dx = r.mean2d[0] - float(px), dy = r.mean2d[1] - float(py);
                    float alpha, G;
                    if (!pair_alpha(r.conic[0], r.conic[1], r.conic[2], r.opacity, dx, dy, g_alpha_clip, alpha, G))
                        continue;
                    const float test_t = Tr * (1.0f - alpha);
                    if (test_t < K<float>::transmittance_floor()) break;
                    const float a_t = alpha * Tr;
                    for (int ch = 0; ch < 3; ++ch) C[ch] = fma_(r.rgb[ch], a_t, C[ch]);
                    Tr = test_t;
                }
                const int64_t pid = int64_t(py) * W + px;
                for (int ch = 0; ch < 3; ++ch)
                    color_concat[3 * vi.pixel_base + ch * P + pid] = C[ch] + Tr * vi.cam.background[ch];
                alpha_concat[vi.pixel_base + pid] = 1.0f - Tr;
                final_T[vi.pixel_base + pid] = Tr;
            }
    }
    return 0;
}

// Images of the last forward at its own precision (double for precision 64): colour per view
// [3,H,W] and alpha per view [H,W], each concatenated over views.
int oracle_images(int precision, double* color, double* alpha) {
    if (precision == 64 || precision == 65 || precision == 66) {
        std::copy(g_state64.color.begin(), g_state64.color.end(), color);
        std::copy(g_state64.alpha.begin(), g_state64.alpha.end(), alpha);
    } else {
        std::copy(g_state32.color.begin(), g_state32.color.end(), color);
        std::copy(g_state32.alpha.begin(), g_state32.alpha.end(), alpha);
    }
    return 0;
}

// Backward of the last forward with the same precision. Gradients wrt raw parameters summed
// over views; double precision writes doubles (pointers are double*).
int oracle_backward(int precision, const pht_gaussians* g, const pht_settings* s, const pht_view_grads* dl,
                    void* means, void* log_scales, void* quats, void* logits, void* sh, void* colors,
                    void* refine, int threads) {
    const int64_t N = g->count;
    if (precision == 66) return -1; // order-matched FP64 renders images only
    if (precision == 64 || precision == 65) {
        std::vector<SplatGradAcc<double>>& acc = g_acc64;
        run_blend_backward<double>(*g, *s, g_state64, dl, acc, threads, precision == 65 ? &g_state32 : nullptr);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t i = 0; i < N; ++i) {
            RowGrad<double> r;
            gaussian_row_grad<double>(*g, *s, g_state64, acc, i, r);
            write_grads<double, double>(*g, *s, i, r, (double*)means, (double*)log_scales, (double*)quats,
                                        (double*)logits, (double*)sh, (double*)colors, (double*)refine);
        }
        return 0;
    }
    std::vector<SplatGradAcc<float>>& acc = g_acc32;
    run_blend_backward<float>(*g, *s, g_state32, dl, acc, threads);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < N; ++i) {
        RowGrad<float> r;
        gaussian_row_grad<float>(*g, *s, g_state32, acc, i, r);
        write_grads<float, float>(*g, *s, i, r, (float*)means, (float*)log_scales, (float*)quats,
                                  (float*)logits, (float*)sh, (float*)colors, (float*)refine);
    }
    return 0;
}

// Per-(gaussian, view) blend gradients of the last backward: out[(g * n_views + v) * 10 + k]
// (d_mean2d.xy, d_conic.xyz, d_opacity, d_color.rgb, refine), zero where not visible.
int oracle_splat_grads(int precision, int64_t n, double* out) {
    if (precision == 64 || precision == 65) export_splat_grads<double>(g_state64, g_acc64, n, out);
    else export_splat_grads<float>(g_state32, g_acc32, n, out);
    return 0;
}

// L1 loss gradient (cuda_ops.cu:886-898): grad = weight / P / 3 * sign(diff); returns the loss.
double oracle_l1_loss(const float* color, const float* target, float* grad, int64_t pixels, float weight) {
    const float inv_p = 1.0f / float(pixels);
    double loss = 0.0;
    for (int64_t c = 0; c < 3 * pixels; ++c) {
        const float d = color[c] - target[c];
        loss += std::fabs(double(d));
        const float sgn = float((d > 0.0f) - (d < 0.0f));
        grad[c] = ((weight * inv_p) / 3.0f) * sgn;
    }
    return loss * double(weight) / (3.0 * double(pixels));
}

// Fused L1 + SSIM of one [3,H,W] view (mask [H,W] or NULL): precision 32 writes float grad,
// 64 writes double grad; returns the loss.
double oracle_l1_ssim(int precision, const float* color, const float* target, const float* mask, int width, int height,
                      float ssim_weight, float photometric_weight, void* grad) {
    if (precision == 64)
        return l1_ssim<double>(color, target, mask, width, height, ssim_weight, photometric_weight,
                               static_cast<double*>(grad));
    return l1_ssim<float>(color, target, mask, width, height, ssim_weight, photometric_weight,
                          static_cast<float*>(grad));
}

// Standalone Adam over rows x row_stride with an active prefix (see pht_adam_step).
void oracle_adam_step(float* p, const float* gr, float* m, float* v, int64_t rows, int64_t stride,
                      int64_t active, float lr, float lr2, int64_t secondary_from, float b1, float b2,
                      float c1, float c2, float eps, float lo, float hi, int prefix) {
    for (int64_t r = 0; r < rows; ++r)
        for (int64_t c = 0; c < active; ++c) {
            const int64_t j = r * stride + c;
            const float lrx = (secondary_from > 0 && c >= secondary_from) ? lr2 : lr;
            adam_update(p[j], m[j], v[j], gr[j], lrx, b1, b2, c1, c2, eps, lo, hi, prefix != 0);
        }
}

// Standalone quantised SH Adam over rows x stride (see pht_sh_adam_quant_step).
void oracle_sh_adam_quant_step(float* p, const float* gr, uint16_t* first, uint8_t* packed, float* bounds,
                               int64_t rows, int stride, int active, float lr, float rest_lr, float b1, float b2,
                               float c1, float c2, float eps, float regularization) {
    const ShqStep q{lr, rest_lr, b1, b2, c1, c2, eps, regularization, active};
    for (int64_t r = 0; r < rows; ++r)
        shq_row(q, stride, p + r * stride, gr ? gr + r * stride : nullptr, first + r * stride, packed + r * stride,
                bounds + 4 * r);
}

// One full training step on the CPU (forward + L1 + backward + fused Adam) in float: the CPU
// baseline. Returns the loss. targets/grad buffers are per view [3,H,W].
// One training step (forward, photometric loss, backward, fused Adam) in float: loss_kind 0 =
// L1 with weight photometric_weight, 1 = fused L1 + SSIM. Returns the loss.
double oracle_train_step(const pht_gaussians* g, const pht_camera* cams, int n_views, const pht_settings* s,
                         pht_view_outputs* outs, const float* const* targets, float* const* grad_bufs,
                         const pht_adam* adam, int threads, int loss_kind, float ssim_weight,
                         float photometric_weight) {
    State<float>& st = g_state32;
    run_forward<float>(*g, cams, n_views, *s, st, outs, threads);
    double loss = 0.0;
    std::vector<pht_view_grads> dl(static_cast<size_t>(n_views));
    for (int v = 0; v < n_views; ++v) {
        const int64_t P = int64_t(cams[v].width) * cams[v].height;
        if (loss_kind == 1)
            loss += l1_ssim<float>(outs[v].color, targets[v], nullptr, cams[v].width, cams[v].height, ssim_weight,
                                   photometric_weight, grad_bufs[v]);
        else
            loss += oracle_l1_loss(outs[v].color, targets[v], grad_bufs[v], P, photometric_weight);
        dl[size_t(v)].grad_color = grad_bufs[v];
        dl[size_t(v)].grad_alpha = nullptr;
    }
    std::vector<SplatGradAcc<float>> acc;
    run_blend_backward<float>(*g, *s, st, dl.data(), acc, threads);
    const int64_t N = g->count;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t i = 0; i < N; ++i) {
        RowGrad<float> r;
        gaussian_row_grad<float>(*g, *s, st, acc, i, r);
        adam_row(*adam, *s, i, r.mean, r.log_scale, r.quat, r.logit, r.sh);
    }
    return loss;
}

// ---------------------------------------------------------------- densification rows
// The refinement's per-row steps (photara/src/splat/cuda_ops.cu:263-302, 1548-1606, 1720-1861;
// igs_selection_cuda.cu:31-70; densification_igs.cpp:439-497), one row at a time over the same
// shared row functions as the kernels. Selection itself is done by the Python oracle by sorting
// the keys (a stable ordering, no radix select).
void oracle_densify_accumulate(const pht_densify_stats* s, const pht_densify_observation* obs, int n_obs, int64_t n) {
    for (int v = 0; v < n_obs; ++v) {
        const int32_t side = std::max<int32_t>(1, std::min(obs[v].width, obs[v].height));
        const float inv = 1.0f / float(side);
        for (int64_t i = 0; i < n; ++i) {
            DensifyRow r = densify_load_row(*s, i);
            densify_accumulate_row(r, obs[v], inv, i);
            densify_store_row(*s, i, r);
        }
    }
}

void oracle_prune_masks(const pht_gaussians* g, int sh_stride, float min_opacity, float max_bounds,
                        const float* center, uint8_t* keep, uint8_t* hard, float* opacities) {
    for (int64_t i = 0; i < g->count; ++i) {
        bool k, h;
        opacities[i] = densify_prune_row(g->means, g->log_scales, g->quaternions, g->opacity_logits, g->sh, sh_stride,
                                         i, min_opacity, max_bounds, center, k, h);
        keep[i] = k ? 1 : 0;
        hard[i] = h ? 1 : 0;
    }
}

// Keys of pht_select_rows (0 = not eligible).
void oracle_select_keys(const float* values, const uint8_t* mask, const uint8_t* exclude, int64_t n, int mode,
                        uint64_t seed, uint64_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        uint64_t key = 0;
        if ((mask == nullptr || mask[i] != 0) && (exclude == nullptr || exclude[i] == 0)) {
            const float v = values[i];
            if (mode == PHT_SELECT_GUMBEL)
                key = gumbel_key(v, seed, uint32_t(i));
            else if (v == v && (mode != PHT_SELECT_LARGEST_POSITIVE || (v > 0.0f && finite_(v))))
                key = select_value_key(v, mode != PHT_SELECT_SMALLEST, uint32_t(i));
        }
        keys[i] = key;
    }
}

void oracle_densify_split(const pht_split_params* p, const int32_t* parents, int64_t count, int64_t child_base,
                          const float* samples, const float* screens) {
    for (int64_t j = 0; j < count; ++j) {
        const int64_t parent = parents[j], child = child_base + j;
        densify_split_row(p->means, p->log_scales, p->opacity_logits, p->quaternions, parent, child, samples + 3 * j,
                          screens ? screens[j] : 0.0f, screens != nullptr, p->mode, p->minimum_opacity,
                          p->split_at_screen_size, p->split_opacity_k);
        if (p->lp_format != PHT_LP_NONE) {
            for (int k = 0; k < 3; ++k) {
                p->log_scales_lp[3 * parent + k] = f32_to_lp(p->log_scales[3 * parent + k], p->lp_format);
                p->log_scales_lp[3 * child + k] = f32_to_lp(p->log_scales[3 * child + k], p->lp_format);
            }
            p->opacity_logits_lp[parent] = f32_to_lp(p->opacity_logits[parent], p->lp_format);
            p->opacity_logits_lp[child] = f32_to_lp(p->opacity_logits[child], p->lp_format);
        }
    }
}

void oracle_adc_decay(float* log_scales, float* logits, uint16_t* ls_lp, uint16_t* logit_lp, int fmt, int64_t n,
                      float opacity_decay, float log_scale_decay) {
    for (int64_t i = 0; i < n; ++i) {
        densify_decay_row(log_scales, logits, i, opacity_decay, log_scale_decay);
        if (fmt != PHT_LP_NONE) {
            for (int k = 0; k < 3; ++k) ls_lp[3 * i + k] = f32_to_lp(log_scales[3 * i + k], fmt);
            logit_lp[i] = f32_to_lp(logits[i], fmt);
        }
    }
}

// pht_igs_prepare + pht_igs_weights over the retained rows.
void oracle_igs_rows(const pht_densify_stats* s, const int32_t* keep_rows, int64_t retained, const float* opacities,
                     float median, int allow_growth, float gthr, float sthr, float* priority, uint8_t* positive,
                     uint32_t* counts, float* replacement, float* growth, float* oversize) {
    counts[0] = counts[1] = counts[2] = 0;
    for (int64_t r = 0; r < retained; ++r) {
        const int64_t src = keep_rows ? keep_rows[r] : r;
        const IgsRow row = igs_row(*s, src, gthr, sthr);
        priority[r] = row.priority;
        positive[r] = row.positive ? 1 : 0;
        counts[0] += row.positive;
        counts[1] += row.growth;
        counts[2] += row.oversized;
        if (replacement) {
            float rw, gw, ow;
            igs_weights_row(*s, src, opacities[src], median, allow_growth != 0, gthr, sthr, rw, gw, ow);
            replacement[r] = rw;
            growth[r] = gw;
            oversize[r] = ow;
        }
    }
}

} // extern "C"

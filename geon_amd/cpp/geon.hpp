// geon.hpp — C++17 face of the MI355X geodesic engine with the public surface
// of the reference crate (Arth-Singh/geon src/lib.rs:18-30): Scalar, Vec4,
// Mat4, the Metric trait and its four metrics, Termination, Camera,
// RenderConfig, RenderStats, render, trace, write_png, plus the curvature,
// sky and integrator functions its public modules export. Header-only, over
// the C ABI of include/geon_hip.h (link libgeon_hip.so).
//
// Differences a caller sees:
//   * the reference panics (camera.rs:53,85; types.rs:39; render.rs:181);
//     these throw geon::Error carrying the ABI status;
//   * render and trace run on the GPU through one process-wide context
//     (device 0, created on first use, calls serialised), where the
//     reference runs rayon over the host's cores;
//   * a metric is one of the four below: user-defined metrics have no GPU
//     code (the reference's generic code stays the path for those).
#pragma once

#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/geon_hip.h"

namespace geon {

// types.rs:7-19
using Scalar = double;
using Vec4 = std::array<Scalar, 4>;
using Mat4 = std::array<std::array<Scalar, 4>, 4>;
using Christoffel = std::array<Mat4, 4>;                  // [μ][α][β]
using Riemann = std::array<std::array<Mat4, 4>, 4>;        // [ρ][σ][μ][ν]
using State = std::array<Scalar, 8>;                       // integrator.rs:12

// A failure the reference reports by panicking, or one of the GPU runtime.
class Error : public std::runtime_error {
public:
    Error(int status, const std::string& where)
        : std::runtime_error(where + ": " + geon_strerror(status) +
                             (geon_last_error()[0] ? std::string(" (") + geon_last_error() + ")"
                                                   : std::string())),
          status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char* where) {
    if (status != GEON_OK) throw Error(status, where);
}

// metric/mod.rs:111-120
enum class Termination { Captured, Escaped, Diverged };

namespace detail {
inline Mat4 mat4(const double* a) {
    Mat4 m;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i][j] = a[4 * i + j];
    return m;
}
inline void flat16(const Mat4& m, double* a) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) a[4 * i + j] = m[i][j];
}
}  // namespace detail

// The Metric trait (metric/mod.rs:19-109). Every method is evaluated on the
// host by the library (geon_metric_*), following the reference's definition
// for that metric; trace, render and the curvature functions run on the GPU.
class Metric {
public:
    virtual ~Metric() = default;
    // The ABI description of this metric (kind, parameters, termination radii).
    virtual geon_metric_t to_c() const = 0;
    virtual const char* name() const = 0;

    Mat4 g(const Vec4& x) const {
        double out[16];
        const geon_metric_t m = to_c();
        check(geon_metric_g(&m, x.data(), out), "Metric::g");
        return detail::mat4(out);
    }
    std::array<Mat4, 4> dg(const Vec4& x) const {
        double out[64];
        const geon_metric_t m = to_c();
        check(geon_metric_dg(&m, x.data(), out), "Metric::dg");
        std::array<Mat4, 4> r;
        for (int a = 0; a < 4; ++a) r[a] = detail::mat4(out + 16 * a);
        return r;
    }
    Mat4 g_inv(const Vec4& x) const {
        double out[16];
        const geon_metric_t m = to_c();
        check(geon_metric_g_inv(&m, x.data(), out), "Metric::g_inv");
        return detail::mat4(out);
    }
    Christoffel christoffel(const Vec4& x) const {
        double out[64];
        const geon_metric_t m = to_c();
        check(geon_metric_christoffel(&m, x.data(), out), "Metric::christoffel");
        Christoffel c;
        for (int mu = 0; mu < 4; ++mu) c[mu] = detail::mat4(out + 16 * mu);
        return c;
    }
    // mod.rs:73-75
    std::array<Scalar, 4> fd_step(const Vec4&) const {
        const Scalar h = to_c().fd_step;
        return {h, h, h, h};
    }
    std::optional<Termination> terminate(const Vec4& x) const {
        int32_t cls = -1;
        const geon_metric_t m = to_c();
        check(geon_metric_terminate(&m, x.data(), &cls), "Metric::terminate");
        if (cls == GEON_CAPTURED) return Termination::Captured;
        if (cls == GEON_ESCAPED) return Termination::Escaped;
        if (cls == GEON_DIVERGED) return Termination::Diverged;
        return std::nullopt;
    }
    std::array<Scalar, 3> asymptotic_direction(const Vec4& x, const Vec4& p) const {
        std::array<Scalar, 3> d;
        const geon_metric_t m = to_c();
        check(geon_metric_asymptotic_direction(&m, x.data(), p.data(), d.data()),
              "Metric::asymptotic_direction");
        return d;
    }
};

namespace detail {
inline geon_metric_t base(int32_t kind) {
    geon_metric_t m;
    check(geon_metric_default(kind, &m), "geon_metric_default");
    return m;
}
}  // namespace detail

// minkowski.rs:8-44 (escape r² > 1e6)
struct Minkowski final : Metric {
    geon_metric_t to_c() const override { return detail::base(GEON_METRIC_MINKOWSKI); }
    const char* name() const override { return "minkowski"; }
};

// schwarzschild.rs:12-126; Schwarzschild::new(mass) is Schwarzschild(mass)
struct Schwarzschild final : Metric {
    Scalar mass;
    Scalar horizon_pad = 1.001;
    Scalar escape_radius = 1000.0;
    explicit Schwarzschild(Scalar mass_) : mass(mass_) {}
    geon_metric_t to_c() const override {
        geon_metric_t m = detail::base(GEON_METRIC_SCHWARZSCHILD);
        m.p[0] = mass;
        m.horizon_pad = horizon_pad;
        m.escape_radius = escape_radius;
        return m;
    }
    const char* name() const override { return "schwarzschild"; }
};

// alcubierre.rs:14-89; Alcubierre::default() is Alcubierre{} (escape r² > 1e6)
struct Alcubierre final : Metric {
    Scalar bubble_radius = 2.0;
    Scalar bubble_thickness = 0.5;
    Scalar ship_velocity = 2.0;
    Scalar ship_x = 0.0;
    geon_metric_t to_c() const override {
        geon_metric_t m = detail::base(GEON_METRIC_ALCUBIERRE);
        m.p[0] = bubble_radius;
        m.p[1] = bubble_thickness;
        m.p[2] = ship_velocity;
        m.p[3] = ship_x;
        return m;
    }
    const char* name() const override { return "alcubierre"; }
};

// morris_thorne.rs:11-78; MorrisThorne::default() is MorrisThorne{} (escape |l| > 1000)
struct MorrisThorne final : Metric {
    Scalar throat_radius = 1.0;
    geon_metric_t to_c() const override {
        geon_metric_t m = detail::base(GEON_METRIC_MORRIS_THORNE);
        m.p[0] = throat_radius;
        return m;
    }
    const char* name() const override { return "morris_thorne"; }
};

// types.rs:22-31 — Σ_i Σ_j g_ij a_i b_j, i-major, left to right
inline Scalar dot4(const Mat4& g, const Vec4& a, const Vec4& b) {
    Scalar s = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) s += g[i][j] * a[i] * b[j];
    return s;
}

// types.rs:34-46 — the inverse by cofactors over the determinant; a zero
// determinant is the reference's "metric became singular" panic
inline Mat4 inv4(const Mat4& m) {
    double a[16], c[16];
    detail::flat16(m, a);
    auto minor3 = [&](int r, int col) {
        int rs[3], cs[3];
        for (int i = 0, k = 0; i < 4; ++i)
            if (i != r) rs[k++] = i;
        for (int j = 0, k = 0; j < 4; ++j)
            if (j != col) cs[k++] = j;
        auto e = [&](int i, int j) { return a[4 * rs[i] + cs[j]]; };
        return e(0, 0) * (e(1, 1) * e(2, 2) - e(1, 2) * e(2, 1)) -
               e(0, 1) * (e(1, 0) * e(2, 2) - e(1, 2) * e(2, 0)) +
               e(0, 2) * (e(1, 0) * e(2, 1) - e(1, 1) * e(2, 0));
    };
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) c[4 * i + j] = ((i + j) % 2 ? -1.0 : 1.0) * minor3(i, j);
    const double det = a[0] * c[0] + a[1] * c[1] + a[2] * c[2] + a[3] * c[3];
    if (det == 0.0) throw Error(GEON_E_SINGULAR, "inv4");
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[i][j] = c[4 * j + i] / det;   // adjugate = cofactorᵀ
    return r;
}

// geodesic.rs:28-36
inline Vec4 pos(const State& y) { return {y[0], y[1], y[2], y[3]}; }
inline Vec4 mom(const State& y) { return {y[4], y[5], y[6], y[7]}; }

// geodesic.rs:38-59 — dx^μ = p^μ, dp^μ = −Σ_αβ (Γ^μ_αβ p^α) p^β (host)
inline State rhs(const Metric& metric, const State& y) {
    const Vec4 p = mom(y);
    const Christoffel G = metric.christoffel(pos(y));
    State out;
    for (int mu = 0; mu < 4; ++mu) {
        out[mu] = p[mu];
        Scalar acc = 0.0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) acc += G[mu][a][b] * p[a] * p[b];
        out[4 + mu] = -acc;
    }
    return out;
}

// integrator.rs:16-40 — classic RK4 for any right-hand side f(y) (host)
template <class F>
State rk4_step(const State& y, Scalar h, F&& f) {
    const State k1 = f(y);
    State t;
    for (int i = 0; i < 8; ++i) t[i] = y[i] + 0.5 * h * k1[i];
    const State k2 = f(t);
    for (int i = 0; i < 8; ++i) t[i] = y[i] + 0.5 * h * k2[i];
    const State k3 = f(t);
    for (int i = 0; i < 8; ++i) t[i] = y[i] + h * k3[i];
    const State k4 = f(t);
    State out;
    for (int i = 0; i < 8; ++i)
        out[i] = y[i] + (h / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    return out;
}

// The process-wide GPU context the GPU-backed functions share.
class Context {
public:
    static Context& global() {
        static Context c;
        return c;
    }
    geon_ctx_t* handle() {
        std::call_once(once_, [&] { status_ = geon_ctx_create(0, nullptr, &ctx_); });
        check(status_, "geon_ctx_create");
        return ctx_;
    }
    std::mutex& mutex() { return mu_; }
    ~Context() {
        if (ctx_) geon_ctx_destroy(ctx_);
    }

private:
    Context() = default;
    std::once_flag once_;
    int status_ = GEON_OK;
    geon_ctx_t* ctx_ = nullptr;
    std::mutex mu_;
};

// geodesic.rs:18-25
struct TraceResult {
    Termination termination;
    State final_state;
    std::size_t steps_taken;
    Scalar null_residual;
};

namespace detail {
inline Termination termination(int32_t cls) {
    if (cls == GEON_CAPTURED) return Termination::Captured;
    if (cls == GEON_ESCAPED) return Termination::Escaped;
    return Termination::Diverged;   // diverged, or out of steps (geodesic.rs:107-116)
}
inline uint32_t steps_u32(std::size_t n) {
    if (n > 0xffffffffu) throw Error(GEON_E_INVALID, "max_steps above 2^32 - 1");
    return (uint32_t)n;
}
}  // namespace detail

// geodesic.rs:62-117, for many rays at once on the GPU
inline std::vector<TraceResult> trace_many(const Metric& metric, const std::vector<State>& y0,
                                           Scalar step, std::size_t max_steps) {
    const std::size_t n = y0.size();
    std::vector<TraceResult> out(n);
    if (n == 0) return out;
    std::vector<double> soa(8 * n), yf(8 * n), resid(n);
    std::vector<int32_t> cls(n);
    std::vector<uint32_t> steps(n);
    for (std::size_t k = 0; k < n; ++k)
        for (int c = 0; c < 8; ++c) soa[c * n + k] = y0[k][c];
    const geon_metric_t m = metric.to_c();
    Context& ctx = Context::global();
    {
        std::lock_guard<std::mutex> lock(ctx.mutex());
        check(geon_trace_rays(ctx.handle(), &m, soa.data(), n, step, detail::steps_u32(max_steps),
                              cls.data(), steps.data(), yf.data(), resid.data(), nullptr),
              "trace");
    }
    for (std::size_t k = 0; k < n; ++k) {
        out[k].termination = detail::termination(cls[k]);
        for (int c = 0; c < 8; ++c) out[k].final_state[c] = yf[c * n + k];
        out[k].steps_taken = steps[k];
        out[k].null_residual = resid[k];
    }
    return out;
}

// geodesic.rs:62-117
inline TraceResult trace(const Metric& metric, const State& y0, Scalar step,
                         std::size_t max_steps) {
    return trace_many(metric, {y0}, step, max_steps)[0];
}

// camera.rs:21-160
struct Camera {
    Vec4 observer;
    Scalar fov_y_rad;
    std::size_t width, height;

    // Camera::new
    Camera(const Vec4& observer_, Scalar fov_y_deg, std::size_t width_, std::size_t height_)
        : observer(observer_), width(width_), height(height_) {
        geon_camera_t c;
        check(geon_camera_new(observer_.data(), fov_y_deg, (uint32_t)width_, (uint32_t)height_, &c),
              "Camera::new");
        fov_y_rad = c.fov_y_rad;
    }
    geon_camera_t to_c() const {
        if (width > 0xffffffffu || height > 0xffffffffu) throw Error(GEON_E_INVALID, "camera size");
        geon_camera_t c;
        std::memcpy(c.observer, observer.data(), sizeof c.observer);
        c.fov_y_rad = fov_y_rad;
        c.width = (uint32_t)width;
        c.height = (uint32_t)height;
        return c;
    }
    Mat4 tetrad(const Metric& metric) const {
        double t[16];
        const geon_metric_t m = metric.to_c();
        check(geon_tetrad(&m, observer.data(), t), "Camera::tetrad");
        return detail::mat4(t);
    }
    State ray_state(const Mat4& tet, std::size_t i, std::size_t j) const {
        return ray_state_sub(tet, i, j, 0.5, 0.5);
    }
    State ray_state_sub(const Mat4& tet, std::size_t i, std::size_t j, Scalar du, Scalar dv) const {
        double t[16];
        detail::flat16(tet, t);
        const geon_camera_t c = to_c();
        State y;
        check(geon_ray_state(&c, t, (uint32_t)i, (uint32_t)j, du, dv, y.data()), "ray_state");
        return y;
    }
};

// render.rs:17-41
struct RenderConfig {
    Scalar step = 0.05;
    std::size_t max_steps = 20000;
    std::array<float, 3> captured_rgb{0.0f, 0.0f, 0.0f};
    std::array<float, 3> diverged_rgb{1.0f, 0.0f, 1.0f};
    std::size_t samples_per_axis = 1;
};

// render.rs:43-51
struct RenderStats {
    std::size_t captured = 0, escaped = 0, diverged = 0;
    Scalar max_null_residual = 0.0;
    double avg_steps = 0.0;
};

// sky.rs:16-120
namespace sky {
// sky.rs:23 — read by render and sample
inline std::atomic<bool> SHOW_STARS{true};
inline std::array<float, 3> sample(const std::array<Scalar, 3>& dir_cart) {
    std::array<float, 3> rgb;
    check(geon_sky_sample(dir_cart.data(), SHOW_STARS.load(std::memory_order_relaxed) ? 1 : 0,
                          rgb.data()),
          "sky::sample");
    return rgb;
}
}  // namespace sky

// render.rs:55-165 — linear RGB (row-major, top-left origin) and statistics
inline std::pair<std::vector<float>, RenderStats> render(const Metric& metric, const Camera& camera,
                                                        const RenderConfig& cfg) {
    geon_render_cfg_t c;
    std::memset(&c, 0, sizeof c);
    c.step = cfg.step;
    c.max_steps = detail::steps_u32(cfg.max_steps);
    c.samples_per_axis = (uint32_t)(cfg.samples_per_axis ? cfg.samples_per_axis : 1);
    std::memcpy(c.captured_rgb, cfg.captured_rgb.data(), sizeof c.captured_rgb);
    std::memcpy(c.diverged_rgb, cfg.diverged_rgb.data(), sizeof c.diverged_rgb);
    c.show_stars = sky::SHOW_STARS.load(std::memory_order_relaxed) ? 1 : 0;
    c.precision = GEON_PREC_FP64;
    const geon_metric_t m = metric.to_c();
    const geon_camera_t cam = camera.to_c();
    std::vector<float> pixels(3 * camera.width * camera.height);
    geon_render_stats_t st;
    Context& ctx = Context::global();
    {
        std::lock_guard<std::mutex> lock(ctx.mutex());
        check(geon_render(ctx.handle(), &m, &cam, &c, pixels.data(), nullptr, nullptr, &st), "render");
    }
    RenderStats s;
    s.captured = st.captured;
    s.escaped = st.escaped;
    s.diverged = st.diverged;
    s.max_null_residual = st.max_null_residual;
    s.avg_steps = st.avg_steps;
    return {std::move(pixels), s};
}

// curvature.rs:48-141 on the GPU (one point; the _at versions take many)
namespace detail {
inline std::vector<double> curvature(const Metric& metric, const std::vector<Vec4>& xs, int32_t what,
                                     std::size_t width) {
    const std::size_t n = xs.size();
    std::vector<double> soa(4 * n), out(width * n);
    if (n == 0) return out;
    for (std::size_t k = 0; k < n; ++k)
        for (int c = 0; c < 4; ++c) soa[c * n + k] = xs[k][c];
    const geon_metric_t m = metric.to_c();
    Context& ctx = Context::global();
    std::lock_guard<std::mutex> lock(ctx.mutex());
    check(geon_curvature(ctx.handle(), &m, soa.data(), n, what, out.data()), "curvature");
    return out;
}
inline Mat4 mat_at(const std::vector<double>& v, std::size_t k) { return mat4(v.data() + 16 * k); }
}  // namespace detail

inline Riemann riemann(const Metric& metric, const Vec4& x) {
    const std::vector<double> v = detail::curvature(metric, {x}, GEON_CURV_RIEMANN, 256);
    Riemann r;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) r[a][b] = detail::mat4(v.data() + 64 * a + 16 * b);
    return r;
}
inline Mat4 ricci(const Metric& metric, const Vec4& x) {
    return detail::mat_at(detail::curvature(metric, {x}, GEON_CURV_RICCI, 16), 0);
}
inline Scalar ricci_scalar(const Metric& metric, const Vec4& x) {
    return detail::curvature(metric, {x}, GEON_CURV_SCALAR, 1)[0];
}
inline Mat4 einstein(const Metric& metric, const Vec4& x) {
    return detail::mat_at(detail::curvature(metric, {x}, GEON_CURV_EINSTEIN, 16), 0);
}
inline Mat4 stress_energy(const Metric& metric, const Vec4& x) {
    return detail::mat_at(detail::curvature(metric, {x}, GEON_CURV_STRESS, 16), 0);
}
inline std::vector<Mat4> stress_energy_at(const Metric& metric, const std::vector<Vec4>& xs) {
    const std::vector<double> v = detail::curvature(metric, xs, GEON_CURV_STRESS, 16);
    std::vector<Mat4> out(xs.size());
    for (std::size_t k = 0; k < xs.size(); ++k) out[k] = detail::mat_at(v, k);
    return out;
}
// curvature.rs:127-136 — T_μν k^μ k^ν
inline Scalar nec_contraction(const Metric& metric, const Vec4& x, const Vec4& k) {
    const geon_metric_t m = metric.to_c();
    const double xs[4] = {x[0], x[1], x[2], x[3]};
    double out = 0.0;
    Context& ctx = Context::global();
    std::lock_guard<std::mutex> lock(ctx.mutex());
    check(geon_nec_contraction(ctx.handle(), &m, xs, 1, k.data(), &out), "nec_contraction");
    return out;
}
// curvature.rs:139-141 — T_μν u^μ u^ν, the same contraction
inline Scalar wec_contraction(const Metric& metric, const Vec4& x, const Vec4& u) {
    return nec_contraction(metric, x, u);
}

// render.rs:169-184 — 8-bit RGB PNG of the linear buffer through the tonemap
// clamp(0,1)^(1/2.2)·255 rounded. The image data is written as stored
// (uncompressed) deflate blocks: the decoded pixels are the reference's,
// the file bytes are not the image crate's.
namespace detail {
inline uint32_t crc32(const uint8_t* p, std::size_t n, uint32_t c = 0xffffffffu) {
    for (std::size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
    }
    return c;
}
inline void put32(std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}
inline void chunk(std::vector<uint8_t>& f, const char* tag, const std::vector<uint8_t>& data) {
    put32(f, (uint32_t)data.size());
    std::vector<uint8_t> td(tag, tag + 4);
    td.insert(td.end(), data.begin(), data.end());
    f.insert(f.end(), td.begin(), td.end());
    put32(f, crc32(td.data(), td.size()) ^ 0xffffffffu);
}
}  // namespace detail

inline void write_png(const std::string& path, const std::vector<float>& pixels, std::size_t width,
                      std::size_t height) {
    if (width == 0 || height == 0 || pixels.size() != 3 * width * height)
        throw Error(GEON_E_INVALID, "write_png: buffer size matches dimensions");
    std::vector<uint8_t> bytes(pixels.size());
    check(geon_tonemap_u8(pixels.data(), pixels.size(), bytes.data()), "geon_tonemap_u8");
    // zlib stream of filter-0 scanlines in stored blocks (≤ 65535 bytes each)
    std::vector<uint8_t> raw;
    raw.reserve(height * (3 * width + 1));
    for (std::size_t j = 0; j < height; ++j) {
        raw.push_back(0);
        raw.insert(raw.end(), bytes.begin() + 3 * width * j, bytes.begin() + 3 * width * (j + 1));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    for (std::size_t off = 0; off < raw.size() || off == 0;) {
        const std::size_t len = std::min<std::size_t>(65535, raw.size() - off);
        z.push_back(off + len == raw.size() ? 1 : 0);
        z.push_back((uint8_t)(len & 0xff));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xff));
        z.push_back((uint8_t)((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (len == 0) break;
    }
    detail::put32(z, (b << 16) | a);
    std::vector<uint8_t> f = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    detail::put32(ihdr, (uint32_t)width);
    detail::put32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    detail::chunk(f, "IHDR", ihdr);
    detail::chunk(f, "IDAT", z);
    detail::chunk(f, "IEND", {});
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw Error(GEON_E_INVALID, "write_png: cannot open " + path);
    const std::size_t w = std::fwrite(f.data(), 1, f.size(), fp);
    const int cl = std::fclose(fp);
    if (w != f.size() || cl != 0) throw Error(GEON_E_INVALID, "write_png: cannot write " + path);
}

}  // namespace geon

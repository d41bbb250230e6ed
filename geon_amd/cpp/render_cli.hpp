// render_cli.hpp — the command line of the reference's render binary
// (src/bin/render.rs:14-67, flags and defaults) and the scene it builds from
// them (render.rs:69-134), for geon-render (render_main.cpp).
//
// Flags (clap's kebab-case long names; "--flag value" or "--flag=value"):
//   --metric <minkowski|schwarzschild|alcubierre|morris-thorne>  [schwarzschild]
//   --width <N> [1024]  --height <N> [1024]  --fov <deg> [60]  --r-obs <r> [20]
//   --step <h> [0.05]  --max-steps <N> [20000]  --samples <N> [1]
//   --no-stars  --out <path> [outputs/render.png]  --bubble-v <v> [0.5]
//   -h/--help, -V/--version
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "geon.hpp"

namespace geon::cli {

enum class MetricKind { Minkowski, Schwarzschild, Alcubierre, MorrisThorne };

struct Args {
    MetricKind metric = MetricKind::Schwarzschild;
    std::size_t width = 1024;
    std::size_t height = 1024;
    double fov = 60.0;           // vertical field of view, degrees
    double r_obs = 20.0;         // observer radial coordinate
    double step = 0.05;          // affine step
    std::size_t max_steps = 20000;
    std::size_t samples = 1;     // √(rays per pixel)
    bool no_stars = false;
    std::string out = "outputs/render.png";
    double bubble_v = 0.5;       // Alcubierre ship velocity
};

// A usage error: clap prints "error: …" and exits with status 2.
struct UsageError {
    std::string message;
};
// --help / --version: print and exit 0.
struct EarlyExit {
    std::string text;
};

inline const char* kind_debug_name(MetricKind k) {   // {:?} of the enum
    switch (k) {
    case MetricKind::Minkowski: return "Minkowski";
    case MetricKind::Schwarzschild: return "Schwarzschild";
    case MetricKind::Alcubierre: return "Alcubierre";
    default: return "MorrisThorne";
    }
}

inline std::string help_text() {
    return "Render a black hole / wormhole / warp bubble.\n\n"
           "Usage: geon-render [OPTIONS]\n\n"
           "Options:\n"
           "      --metric <METRIC>        [default: schwarzschild] [possible values: minkowski, "
           "schwarzschild, alcubierre, morris-thorne]\n"
           "      --width <WIDTH>          [default: 1024]\n"
           "      --height <HEIGHT>        [default: 1024]\n"
           "      --fov <FOV>              Field of view, degrees, vertical [default: 60]\n"
           "      --r-obs <R_OBS>          Observer radial coordinate (units of M for "
           "Schwarzschild) [default: 20]\n"
           "      --step <STEP>            Integrator step size (affine parameter) [default: 0.05]\n"
           "      --max-steps <MAX_STEPS>  Maximum integration steps per ray [default: 20000]\n"
           "      --samples <SAMPLES>      √(rays per pixel). 1=none, 2=4×SS, 3=9×, 4=16× "
           "[default: 1]\n"
           "      --no-stars               Disable the procedural starfield\n"
           "      --out <OUT>              Output PNG path [default: outputs/render.png]\n"
           "      --bubble-v <BUBBLE_V>    Alcubierre bubble velocity (in units of c) "
           "[default: 0.5]\n"
           "  -h, --help                   Print help\n"
           "  -V, --version                Print version\n";
}

namespace detail {
inline double parse_f64(const std::string& flag, const std::string& v) {
    char* end = nullptr;
    errno = 0;
    const double x = std::strtod(v.c_str(), &end);
    if (v.empty() || end != v.c_str() + v.size() || errno == ERANGE)
        throw UsageError{"invalid value '" + v + "' for '" + flag + " <" + flag.substr(2) +
                         ">': invalid float literal"};
    return x;
}
inline std::size_t parse_usize(const std::string& flag, const std::string& v) {
    if (v.empty() || v.find_first_not_of("0123456789") != std::string::npos)
        throw UsageError{"invalid value '" + v + "' for '" + flag + "': invalid digit found in string"};
    errno = 0;
    char* end = nullptr;
    const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
    if (errno == ERANGE) throw UsageError{"invalid value '" + v + "' for '" + flag + "': too large"};
    return (std::size_t)x;
}
inline MetricKind parse_kind(const std::string& v) {
    if (v == "minkowski") return MetricKind::Minkowski;
    if (v == "schwarzschild") return MetricKind::Schwarzschild;
    if (v == "alcubierre") return MetricKind::Alcubierre;
    if (v == "morris-thorne") return MetricKind::MorrisThorne;
    throw UsageError{"invalid value '" + v + "' for '--metric <METRIC>'\n  [possible values: "
                     "minkowski, schwarzschild, alcubierre, morris-thorne]"};
}
}  // namespace detail

inline Args parse_args(const std::vector<std::string>& argv) {
    Args a;
    for (std::size_t i = 0; i < argv.size(); ++i) {
        std::string flag = argv[i], value;
        bool has_value = false;
        const std::size_t eq = flag.find('=');
        if (flag.rfind("--", 0) == 0 && eq != std::string::npos) {
            value = flag.substr(eq + 1);
            flag = flag.substr(0, eq);
            has_value = true;
        }
        if (flag == "-h" || flag == "--help") throw EarlyExit{help_text()};
        if (flag == "-V" || flag == "--version") throw EarlyExit{"geon-render 0.1.0\n"};
        if (flag == "--no-stars") {
            if (has_value) throw UsageError{"unexpected value '" + value + "' for '--no-stars'"};
            a.no_stars = true;
            continue;
        }
        static const char* kValued[] = {"--metric", "--width",   "--height",  "--fov",
                                        "--r-obs",  "--step",    "--max-steps", "--samples",
                                        "--out",    "--bubble-v"};
        bool known = false;
        for (const char* k : kValued) known = known || flag == k;
        if (!known) throw UsageError{"unexpected argument '" + argv[i] + "' found"};
        if (!has_value) {
            if (i + 1 >= argv.size())
                throw UsageError{"a value is required for '" + flag + "' but none was supplied"};
            value = argv[++i];
        }
        if (flag == "--metric") a.metric = detail::parse_kind(value);
        else if (flag == "--width") a.width = detail::parse_usize(flag, value);
        else if (flag == "--height") a.height = detail::parse_usize(flag, value);
        else if (flag == "--fov") a.fov = detail::parse_f64(flag, value);
        else if (flag == "--r-obs") a.r_obs = detail::parse_f64(flag, value);
        else if (flag == "--step") a.step = detail::parse_f64(flag, value);
        else if (flag == "--max-steps") a.max_steps = detail::parse_usize(flag, value);
        else if (flag == "--samples") a.samples = detail::parse_usize(flag, value);
        else if (flag == "--out") a.out = value;
        else a.bubble_v = detail::parse_f64(flag, value);
    }
    return a;
}

// The scene of render.rs:69-134: the metric (Schwarzschild::new(1.0),
// Alcubierre { ship_velocity: bubble_v, ..default }, MorrisThorne::default,
// Minkowski), the observer ([0, r_obs, π/2, 0] in the spherical charts,
// [0, r_obs, 0, 0] in the Cartesian ones), the camera and RenderConfig
// { step, max_steps, samples_per_axis, ..default }.
struct Scene {
    std::unique_ptr<Metric> metric;
    Camera camera;
    RenderConfig cfg;
    bool show_stars;
};

inline Scene scene(const Args& a) {
    const double half_pi = 1.5707963267948966;   // std::f64::consts::FRAC_PI_2
    std::unique_ptr<Metric> m;
    Vec4 obs{0.0, a.r_obs, 0.0, 0.0};
    switch (a.metric) {
    case MetricKind::Minkowski:
        m = std::make_unique<Minkowski>();
        break;
    case MetricKind::Schwarzschild:
        m = std::make_unique<Schwarzschild>(1.0);
        obs[2] = half_pi;
        break;
    case MetricKind::Alcubierre: {
        auto al = std::make_unique<Alcubierre>();
        al->ship_velocity = a.bubble_v;
        m = std::move(al);
        break;
    }
    default:
        m = std::make_unique<MorrisThorne>();
        obs[2] = half_pi;
        break;
    }
    RenderConfig cfg;
    cfg.step = a.step;
    cfg.max_steps = a.max_steps;
    cfg.samples_per_axis = a.samples;
    return Scene{std::move(m), Camera(obs, a.fov, a.width, a.height), cfg, !a.no_stars};
}

// f64 Display of the reference's messages: the shortest decimal that reads
// back to the same double, in positional notation, "60" for 60.0.
inline std::string rust_display(double x) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    if (x == 0.0) return std::signbit(x) ? "-0" : "0";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*e", prec - 1, x);
        if (std::strtod(buf, nullptr) == x) break;
    }
    // digits and decimal exponent of the shortest form
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    const std::size_t e = s.find('e');
    const int exp10 = std::atoi(s.c_str() + e + 1);
    std::string digits;
    for (std::size_t i = 0; i < e; ++i)
        if (s[i] != '.') digits += s[i];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    std::string out;
    const int point = exp10 + 1;   // digits before the decimal point
    if (point <= 0) {
        out = "0." + std::string((std::size_t)-point, '0') + digits;
    } else if ((std::size_t)point >= digits.size()) {
        out = digits + std::string((std::size_t)point - digits.size(), '0');
    } else {
        out = digits.substr(0, (std::size_t)point) + "." + digits.substr((std::size_t)point);
    }
    return (neg ? "-" : "") + out;
}

// {:.prec e} of the reference's messages: d.ddde±x with no exponent padding
// and no '+' ("6.840e-13", "1.000e0").
inline std::string rust_lower_exp(double x, int prec) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, x);
    std::string s(buf);
    const std::size_t e = s.find('e');
    const int exp10 = std::atoi(s.c_str() + e + 1);
    return s.substr(0, e) + "e" + std::to_string(exp10);
}

// {:.1} of a f64
inline std::string rust_fixed(double x, int prec) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.*f", prec, x);
    return buf;
}

// render.rs:80-83 and :136-139
inline std::string banner(const Args& a) {
    return "rendering " + std::string(kind_debug_name(a.metric)) + " @ " + std::to_string(a.width) +
           "x" + std::to_string(a.height) + ", fov " + rust_display(a.fov) + "°, r_obs " +
           rust_display(a.r_obs) + ", step " + rust_display(a.step);
}
inline std::string summary(const RenderStats& s) {
    return "done. captured=" + std::to_string(s.captured) + ", escaped=" + std::to_string(s.escaped) +
           ", diverged=" + std::to_string(s.diverged) + ", avg_steps=" + rust_fixed(s.avg_steps, 1) +
           ", max_null_residual=" + rust_lower_exp(s.max_null_residual, 3);
}

}  // namespace geon::cli

// api_driver.cpp — test driver of the C++ face (geon_amd/cpp/geon.hpp) and of
// the CLI's flag mapping (geon_amd/cpp/render_cli.hpp). Prints JSON.
//
//   api_driver args [flags...]   the parsed flags and the scene they build
//   api_driver format v...       rust_display / {:.3e} / {:.1} of each value
//   api_driver host              Metric methods, dot4, inv4, rhs, rk4_step on
//                                the host at fixed points
//   api_driver gpu               trace_many and render on device 0
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../geon_amd/cpp/render_cli.hpp"

using namespace geon;

// a double that JSON reads back as a float, sign of zero included ("-0.0")
static std::string num(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    std::string s(buf);
    if (s.find_first_of(".eni") == std::string::npos) s += ".0";
    return s;
}

static void arr(const double* v, int n) {
    std::printf("[");
    for (int i = 0; i < n; ++i) std::printf("%s%s", i ? ", " : "", num(v[i]).c_str());
    std::printf("]");
}

static int args_mode(const std::vector<std::string>& argv) {
    try {
        const cli::Args a = cli::parse_args(argv);
        const cli::Scene sc = cli::scene(a);
        const geon_metric_t m = sc.metric->to_c();
        std::printf("{\"kind\": %d, \"name\": \"%s\", \"p\": ", m.kind, sc.metric->name());
        arr(m.p, 6);
        std::printf(", \"horizon_pad\": %.17g, \"escape_radius\": %.17g, \"fd_step\": %.17g",
                    m.horizon_pad, m.escape_radius, m.fd_step);
        std::printf(", \"observer\": ");
        arr(sc.camera.observer.data(), 4);
        std::printf(", \"fov_y_rad\": %.17g, \"width\": %zu, \"height\": %zu", sc.camera.fov_y_rad,
                    sc.camera.width, sc.camera.height);
        std::printf(", \"step\": %.17g, \"max_steps\": %zu, \"samples_per_axis\": %zu",
                    sc.cfg.step, sc.cfg.max_steps, sc.cfg.samples_per_axis);
        std::printf(", \"captured_rgb\": [%g, %g, %g], \"diverged_rgb\": [%g, %g, %g]",
                    sc.cfg.captured_rgb[0], sc.cfg.captured_rgb[1], sc.cfg.captured_rgb[2],
                    sc.cfg.diverged_rgb[0], sc.cfg.diverged_rgb[1], sc.cfg.diverged_rgb[2]);
        std::printf(", \"show_stars\": %s, \"out\": \"%s\", \"banner\": \"%s\"}\n",
                    sc.show_stars ? "true" : "false", a.out.c_str(), cli::banner(a).c_str());
        return 0;
    } catch (const cli::UsageError& e) {
        std::printf("{\"error\": \"usage\"}\n");
        std::fprintf(stderr, "%s\n", e.message.c_str());
        return 2;
    } catch (const cli::EarlyExit&) {
        std::printf("{\"early_exit\": true}\n");
        return 0;
    }
}

static int format_mode(const std::vector<std::string>& v) {
    std::printf("[");
    for (std::size_t i = 0; i < v.size(); ++i) {
        const double x = std::strtod(v[i].c_str(), nullptr);
        std::printf("%s[\"%s\", \"%s\", \"%s\"]", i ? ", " : "", cli::rust_display(x).c_str(),
                    cli::rust_lower_exp(x, 3).c_str(), cli::rust_fixed(x, 1).c_str());
    }
    std::printf("]\n");
    return 0;
}

static int host_mode() {
    Schwarzschild s(1.0);
    MorrisThorne mt;
    Alcubierre al;
    al.bubble_thickness = 0.125;
    al.ship_velocity = 0.5;
    const Vec4 xs{0.0, 7.0, 1.1, 0.3}, xm{0.0, 1.7, 0.9, 0.2}, xa{0.0, 1.9, 0.4, -0.3};
    std::printf("{");
    const Metric* ms[3] = {&s, &mt, &al};
    const Vec4* xp[3] = {&xs, &xm, &xa};
    const char* names[3] = {"schwarzschild", "morris_thorne", "alcubierre"};
    for (int k = 0; k < 3; ++k) {
        const Mat4 g = ms[k]->g(*xp[k]), gi = ms[k]->g_inv(*xp[k]);
        const Christoffel c = ms[k]->christoffel(*xp[k]);
        std::printf("%s\"%s\": {\"g\": ", k ? ", " : "", names[k]);
        arr(&g[0][0], 16);
        std::printf(", \"g_inv\": ");
        arr(&gi[0][0], 16);
        std::printf(", \"gamma\": ");
        arr(&c[0][0][0], 64);
        const Mat4 inv = inv4(g);
        std::printf(", \"inv4\": ");
        arr(&inv[0][0], 16);
        const Vec4 p{1.3, -0.4, 0.02, 0.05};
        std::printf(", \"dot4\": %.17g", dot4(g, p, p));
        State y{0.0, (*xp[k])[1], (*xp[k])[2], (*xp[k])[3], 1.3, -0.4, 0.02, 0.05};
        const State r = rhs(*ms[k], y);
        std::printf(", \"rhs\": ");
        arr(r.data(), 8);
        const State y1 = rk4_step(y, 0.05, [&](const State& q) { return rhs(*ms[k], q); });
        std::printf(", \"rk4\": ");
        arr(y1.data(), 8);
        const auto t = ms[k]->terminate(*xp[k]);
        std::printf(", \"terminate\": %d, \"fd_step\": %.17g}", t ? (int)*t : -1,
                    ms[k]->fd_step(*xp[k])[0]);
    }
    bool threw = false;
    try {
        Mat4 z{};
        inv4(z);
    } catch (const Error& e) {
        threw = e.status() == GEON_E_SINGULAR;
    }
    Camera cam({0.0, 30.0, 1.5707963267948966, 0.0}, 60.0, 16, 8);
    const Mat4 tet = cam.tetrad(s);
    const State y = cam.ray_state(tet, 3, 2);
    std::printf(", \"singular_throws\": %s, \"ray_state\": ", threw ? "true" : "false");
    arr(y.data(), 8);
    std::printf("}\n");
    return 0;
}

static int gpu_mode() {
    Schwarzschild s(1.0);
    s.horizon_pad = 1.005;
    s.escape_radius = 100.0;
    Camera cam({0.0, 30.0, 1.5707963267948966, 0.0}, 60.0, 32, 24);
    const Mat4 tet = cam.tetrad(s);
    std::vector<State> y0;
    for (std::size_t j = 0; j < cam.height; ++j)
        for (std::size_t i = 0; i < cam.width; ++i) y0.push_back(cam.ray_state(tet, i, j));
    const std::vector<TraceResult> tr = trace_many(s, y0, 0.05, 4000);
    RenderConfig cfg;
    cfg.max_steps = 4000;
    const auto [pixels, st] = render(s, cam, cfg);
    const TraceResult one = trace(s, y0[5], 0.05, 4000);
    std::printf("{\"cls\": [");
    for (std::size_t k = 0; k < tr.size(); ++k) {
        const int c = tr[k].termination == Termination::Captured ? 1
                      : tr[k].termination == Termination::Escaped ? 2 : 0;
        std::printf("%s%d", k ? ", " : "", c);
    }
    std::printf("], \"steps\": [");
    for (std::size_t k = 0; k < tr.size(); ++k) std::printf("%s%zu", k ? ", " : "", tr[k].steps_taken);
    std::printf("], \"final\": [");
    for (std::size_t k = 0; k < tr.size(); ++k) {
        std::printf("%s", k ? ", " : "");
        arr(tr[k].final_state.data(), 8);
    }
    std::printf("], \"one_steps\": %zu, \"stats\": {\"captured\": %zu, \"escaped\": %zu, "
                "\"diverged\": %zu, \"max_null_residual\": %.17g, \"avg_steps\": %.17g}",
                one.steps_taken, st.captured, st.escaped, st.diverged, st.max_null_residual,
                st.avg_steps);
    std::printf(", \"pixels\": ");
    std::vector<double> px(pixels.begin(), pixels.end());
    arr(px.data(), (int)px.size());
    MorrisThorne mt;
    const Mat4 T = stress_energy(mt, {0.0, 0.0, 1.5707963267948966, 0.0});
    std::printf(", \"mt_T_tt\": %.17g, \"mt_nec\": %.17g}\n", T[0][0],
                nec_contraction(mt, {0.0, 0.0, 1.5707963267948966, 0.0}, {1.0, 1.0, 0.0, 0.0}));
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    const std::vector<std::string> rest(argv + 2, argv + argc);
    try {
        if (mode == "args") return args_mode(rest);
        if (mode == "format") return format_mode(rest);
        if (mode == "host") return host_mode();
        if (mode == "gpu") return gpu_mode();
    } catch (const Error& e) {
        std::fprintf(stderr, "geon::Error: %s\n", e.what());
        return 1;
    }
    return 2;
}

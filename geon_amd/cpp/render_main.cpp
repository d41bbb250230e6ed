// geon-render — the reference's render binary (src/bin/render.rs) on the
// MI355X: the same flags and defaults, the same scene per metric, the same
// messages; the frame is traced on the GPU and written as a PNG.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <string>
#include <vector>

#include "render_cli.hpp"

namespace {

// std::fs::create_dir_all of the output's parent directory
bool create_dirs(const std::string& path) {
    const std::size_t k = path.find_last_of('/');
    if (k == std::string::npos || k == 0) return true;
    const std::string dir = path.substr(0, k);
    std::string cur;
    for (std::size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
        }
        if (i < dir.size()) cur += dir[i];
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    using namespace geon;
    cli::Args args;
    try {
        args = cli::parse_args(std::vector<std::string>(argv + 1, argv + argc));
    } catch (const cli::EarlyExit& e) {
        std::fputs(e.text.c_str(), stdout);
        return 0;
    } catch (const cli::UsageError& e) {
        std::fprintf(stderr, "error: %s\n\nFor more information, try '--help'.\n", e.message.c_str());
        return 2;
    }
    try {
        if (args.no_stars) sky::SHOW_STARS.store(false, std::memory_order_relaxed);
        cli::Scene sc = cli::scene(args);
        if (!create_dirs(args.out)) {
            std::fprintf(stderr, "Error: cannot create the directory of %s\n", args.out.c_str());
            return 1;
        }
        std::printf("%s\n", cli::banner(args).c_str());
        std::fflush(stdout);
        auto [pixels, stats] = render(*sc.metric, sc.camera, sc.cfg);
        write_png(args.out, pixels, args.width, args.height);
        std::printf("%s\n", cli::summary(stats).c_str());
        std::printf("wrote %s\n", args.out.c_str());
    } catch (const Error& e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// Command-line harness of the SpGEMM apps, with the reference's interface
// (src/apps/common.hpp:22-208):
//
//   <app> inputA inputB [--loglevel L] validate
//   <app> inputA inputB [--loglevel L] save <output>
//   <app> inputA inputB [--loglevel L] benchmark [--runs N] [--warmups N] [--pause MS]
//
// benchmark prints "A: / B: / NIP: / C:" and one "Run %2d: <ns>ns | <flops> FLOPS"
// line per run, the format tasks.py parses; extra statistics follow on a
// "Stats:" line.  validate compares with a host Gustavson product (the
// reference used cuSPARSE ALG2 and csr.cuh:218-239's equality; here the
// structure must match bit-exactly and the values within 1e-12 / 1e-5 of
// the product after |A||B| scaling, and the reference's rule is reported
// beside it) and prints "Validation succeeded" / "Validation failed"; the
// exit code is 0 either
// way, as in the reference.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "gpuspgemm.hpp"
#include "host_spgemm.hpp"

namespace app {

enum class Level { trace, debug, info, warn, error, critical, off };
inline Level g_level = Level::info;

inline bool parse_level(const std::string& s, Level& out) {
    static const char* names[] = {"trace", "debug", "info", "warn", "error", "critical", "off"};
    for (int i = 0; i < 7; ++i)
        if (s == names[i]) {
            out = static_cast<Level>(i);
            return true;
        }
    return false;
}

template <typename... Args>
void log(Level lv, const char* fmt, Args... args) {
    if (lv < g_level || g_level == Level::off) return;
    std::fprintf(stderr, "[%s] ", lv <= Level::debug ? "debug" : lv == Level::info ? "info" : "warning");
    std::fprintf(stderr, fmt, args...);
    std::fprintf(stderr, "\n");
}

struct Options {
    std::string input_a, input_b, command, output;
    int runs = 1, warmups = 0, pause = 50;
};

inline int usage(const char* prog, const std::string& err) {
    if (!err.empty()) std::fprintf(stderr, "%s\n", err.c_str());
    std::fprintf(stderr,
                 "Usage: %s inputA inputB [--loglevel trace|debug|info|warn|error|critical|off] "
                 "{validate | save <output> | benchmark [--runs N] [--warmups N] [--pause MS]}\n",
                 prog);
    return err.empty() ? 0 : 2;
}

// Returns -1 to continue, else an exit code.
inline int parse(int argc, char** argv, Options& o) {
    std::vector<std::string> pos;
    auto value = [&](int& i, const std::string& a, const std::string& name, std::string& out) -> bool {
        if (a.rfind(name + "=", 0) == 0) {
            out = a.substr(name.size() + 1);
            return true;
        }
        if (a == name && i + 1 < argc) {
            out = argv[++i];
            return true;
        }
        return false;
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        std::string v;
        if (a == "-h" || a == "--help") return usage(argv[0], "");
        if (a.rfind("--loglevel", 0) == 0) {
            if (!value(i, a, "--loglevel", v) || !parse_level(v, g_level))
                return usage(argv[0], "--loglevel: value not in {trace,debug,info,warn,error,critical,off}");
            continue;
        }
        if (a.rfind("--runs", 0) == 0 || a.rfind("--warmups", 0) == 0 || a.rfind("--pause", 0) == 0) {
            const std::string name = a.substr(0, a.find('='));
            if (!value(i, a, name, v)) return usage(argv[0], name + ": missing value");
            char* end = nullptr;
            const long x = std::strtol(v.c_str(), &end, 10);
            if (!end || *end || x < 0) return usage(argv[0], name + ": value must be a non-negative number");
            (name == "--runs" ? o.runs : name == "--warmups" ? o.warmups : o.pause) = static_cast<int>(x);
            continue;
        }
        if (!a.empty() && a[0] == '-') return usage(argv[0], "unknown option " + a);
        pos.push_back(a);
    }
    if (pos.size() < 3) return usage(argv[0], "inputA, inputB and exactly one subcommand are required");
    o.input_a = pos[0];
    o.input_b = pos[1];
    o.command = pos[2];
    for (const auto* f : {&o.input_a, &o.input_b})
        if (!std::filesystem::is_regular_file(*f)) return usage(argv[0], "File does not exist: " + *f);
    if (o.command == "save") {
        if (pos.size() != 4) return usage(argv[0], "save: output is required");
        o.output = pos[3];
    } else if (o.command == "validate" || o.command == "benchmark") {
        if (pos.size() != 3) return usage(argv[0], "unexpected extra arguments");
    } else {
        return usage(argv[0], "unknown subcommand " + o.command);
    }
    return -1;
}

template <typename T, typename Run, typename Measure>
int run_app(const char* name, Run run, Measure measure, int argc, char** argv) {
    using Clock = std::chrono::steady_clock;
    Options o;
    if (const int rc = parse(argc, argv, o); rc >= 0) return rc;
    (void)name;
    // the library's logs follow the app's level, as the reference's spdlog level does
    spg_set_log_level(static_cast<int32_t>(g_level));
    try {
        const auto h_a = utils::HostCSR<T>::load_from_filename(o.input_a);
        const auto h_b = utils::HostCSR<T>::load_from_filename(o.input_b);
        if (h_a.n != h_b.m) {
            std::fprintf(stderr, "Matrix A columns (%d) must match matrix B rows (%d)\n", h_a.n, h_b.m);
            return EXIT_FAILURE;
        }
        if (o.command == "validate") {
            const auto h_c = [&] {
                const auto d_a = h_a.template to<utils::Location::Device>();
                const auto d_b = h_b.template to<utils::Location::Device>();
                return run(d_a, d_b);
            }();
            // Structure bit-exact and every value within 1e-12 (FP64) / 1e-5
            // (FP32) of the host product after scaling by (|A||B|)_ij; the
            // reference's own rule (operator==, atol 1e-8 / rtol 1e-5) is
            // reported beside it: in FP32 it fails on cancelling entries.
            const auto h_ref = host::gustavson(h_a, h_b);
            const auto err = host::errors(h_c, h_ref, host::abs_product(h_a, h_b));
            const double tol = sizeof(T) == 8 ? 1e-12 : 1e-5;
            if (err.structure && err.scaled <= tol) {
                std::printf("Validation succeeded\n");
            } else {
                std::printf("Validation failed\n");
                std::printf("Matrix C: %d x %d (%d non-zero elements)\n", h_c.m, h_c.n, h_c.nnz);
                std::printf("Matrix C (host reference): %d x %d (%d non-zero elements)\n", h_ref.m, h_ref.n,
                            h_ref.nnz);
            }
            std::printf("Max error: elementwise relative %.3e, |A||B|-scaled %.3e (tolerance %.0e), structure %s, "
                        "reference rule (atol 1e-8, rtol 1e-5) %s\n",
                        err.elem, err.scaled, tol, err.structure ? "identical" : "different",
                        h_c == h_ref ? "holds" : "fails");
        } else if (o.command == "save") {
            std::printf("Saving result to %s\n", o.output.c_str());
            const auto h_c = [&] {
                const auto d_a = h_a.template to<utils::Location::Device>();
                const auto d_b = h_b.template to<utils::Location::Device>();
                return run(d_a, d_b);
            }();
            h_c.save_to_filename(o.output);
        } else {
            const auto pause = std::chrono::milliseconds(o.pause);
            const auto d_a = h_a.template to<utils::Location::Device>();
            const auto d_b = h_b.template to<utils::Location::Device>();
            for (int i = 0; i < o.warmups; ++i) {
                std::this_thread::sleep_for(pause);
                measure(d_a, d_b);
            }
            std::vector<Clock::duration> times(static_cast<std::size_t>(o.runs));
            for (int i = 0; i < o.runs; ++i) {
                std::this_thread::sleep_for(pause);
                times[static_cast<std::size_t>(i)] = measure(d_a, d_b);
            }
            const auto nip = utils::get_nip(d_a, d_b);
            const double flop = 2.0 * static_cast<double>(nip);
            const auto h_c = run(d_a, d_b);
            std::printf("A: %d x %d (%d non-zero elements)\n", h_a.m, h_a.n, h_a.nnz);
            std::printf("B: %d x %d (%d non-zero elements)\n", h_b.m, h_b.n, h_b.nnz);
            std::printf("NIP: %lld\n", static_cast<long long>(nip));
            std::printf("C: %d x %d (%d non-zero elements)\n", h_c.m, h_c.n, h_c.nnz);
            std::vector<double> secs;
            for (int i = 0; i < o.runs; ++i) {
                const long long ns =
                    std::chrono::duration_cast<std::chrono::nanoseconds>(times[static_cast<std::size_t>(i)]).count();
                const double s = std::chrono::duration<double>(times[static_cast<std::size_t>(i)]).count();
                secs.push_back(s);
                std::printf("Run %2d: %lldns | %.3f FLOPS\n", i + 1, ns, flop / s);
            }
            if (!secs.empty()) {
                std::vector<double> sorted = secs;
                std::sort(sorted.begin(), sorted.end());
                double mean = 0;
                for (double s : secs) mean += s;
                mean /= static_cast<double>(secs.size());
                const auto pct = [&](double q) {
                    return sorted[std::min(sorted.size() - 1, static_cast<std::size_t>(q * (sorted.size() - 1) + 0.5))];
                };
                const double bytes = 4.0 * (h_a.m + 1) + (4.0 + sizeof(T)) * h_a.nnz + 4.0 * (h_b.m + 1) +
                                     (4.0 + sizeof(T)) * h_b.nnz + 4.0 * (h_c.m + 1) + (4.0 + sizeof(T)) * h_c.nnz;
                std::printf("Stats: mean_ms=%.4f p50_ms=%.4f p99_ms=%.4f gflops=%.3f hbm_gbps=%.1f "
                            "hbm_roofline_frac=%.4f host_cores=%u\n",
                            mean * 1e3, pct(0.5) * 1e3, pct(0.99) * 1e3, flop / mean / 1e9, bytes / mean / 1e9,
                            bytes / mean / 8.0e12, std::thread::hardware_concurrency());
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}

}  // namespace app

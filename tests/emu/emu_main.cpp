// Driver of the CPU emulation build: C = A . B through the library's C ABI
// (kernels emulated on the host), compared with the host Gustavson product.
//   emu_spgemm A.csr B.csr [--f32] [--sym-table FILE] [--num-table FILE] [--repeat N]
// Exit status 0 when the structure is identical and values agree to 1e-12
// (FP64) / 1e-5 (FP32) after |A||B| scaling.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "gpuspgemm.hpp"
#include "host_spgemm.hpp"

static std::string slurp(const std::string& f) {
    std::ifstream in(f);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

template <typename T>
static int run(const std::string& fa, const std::string& fb, int repeat) {
    const auto ha = utils::HostCSR<T>::load_from_filename(fa);
    const auto hb = utils::HostCSR<T>::load_from_filename(fb);
    const auto da = ha.template to<utils::Location::Device>();
    const auto db = hb.template to<utils::Location::Device>();
    const auto ref = host::gustavson(ha, hb);
    const auto absref = host::abs_product(ha, hb);
    int bad = 0;
    for (int r = 0; r < repeat; ++r) {
        const auto hc = proposal::proposal(da, db).template to<utils::Location::Host>();
        const auto e = host::errors(hc, ref, absref);
        const double tol = sizeof(T) == 8 ? 1e-12 : 1e-5;
        std::printf("run %d: nnz %d (ref %d) structure %s scaled_err %.3e\n", r, hc.nnz, ref.nnz,
                    e.structure ? "identical" : "DIFFERENT", e.scaled);
        if (!e.structure || e.scaled > tol) bad = 1;
    }
    return bad;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s A.csr B.csr [--f32] [--sym-table F] [--num-table F] [--repeat N]\n", argv[0]);
        return 2;
    }
    bool f32 = false;
    int repeat = 1;
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--f32") {
            f32 = true;
        } else if ((a == "--sym-table" || a == "--num-table") && i + 1 < argc) {
            const std::string t = slurp(argv[++i]);
            utils::handle_spg_error(spg_set_tuning(a == "--sym-table" ? 0 : 1, t.c_str()), "set_tuning");
        } else if (a == "--repeat" && i + 1 < argc) {
            repeat = std::atoi(argv[++i]);
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        return f32 ? run<float>(argv[1], argv[2], repeat) : run<double>(argv[1], argv[2], repeat);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

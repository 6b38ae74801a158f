// The apps' validate checker (gpuspgemm_amd/apps/host_spgemm.hpp) as a
// command, so that a test can compare it with the oracle:
//   host_product A.csr B.csr C.csr      writes C = A.B (FP64)
#include <cstdio>
#include <exception>

#include "host_spgemm.hpp"

// The CSR header names three library entry points (its error helper, and
// the release of library-owned device arrays); this command links no
// library, and host CSRs reach none of them.
extern "C" const char* spg_last_error(void) { return ""; }
extern "C" const char* spg_status_str(spg_status) { return ""; }
extern "C" spg_status spg_csr_free(spg_csr*) { return SPG_OK; }

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: host_product A.csr B.csr C.csr\n");
        return 2;
    }
    try {
        const auto A = utils::HostCSR<double>::load_from_filename(argv[1]);
        const auto B = utils::HostCSR<double>::load_from_filename(argv[2]);
        host::gustavson(A, B).save_to_filename(argv[3]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}

// Driver of the CPU emulation build: C = A . B through the library's C ABI
// (kernels emulated on the host), compared with the host Gustavson product.
//   emu_spgemm A.csr B.csr [--f32] [--sym-table FILE] [--num-table FILE] [--repeat N]
//              [--graphs] [--serve N BUDGET_BYTES MAX_BATCH] [--noalloc]
// Exit status 0 when the structure is identical and values agree to 1e-12
// (FP64) / 1e-5 (FP32) after |A||B| scaling.  --serve submits the product N
// times to a server (spg_server_*) and checks every result; a budget below
// one output must fail every request with SPG_ERR_OUT_OF_MEMORY.
// --graphs turns graph replay on; repeat r then multiplies a fresh device
// copy of A with its values scaled by r + 1 (new addresses, new values,
// same launches), and the graph counters are printed.  --noalloc runs the
// products through spg_multiply_noalloc into caller arrays.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gpuspgemm.hpp"
#include "host_spgemm.hpp"

static std::string slurp(const std::string& f) {
    std::ifstream in(f);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

template <typename T>
static int run(const std::string& fa, const std::string& fb, int repeat, bool graphs) {
    const auto ha = utils::HostCSR<T>::load_from_filename(fa);
    const auto hb = utils::HostCSR<T>::load_from_filename(fb);
    const auto da = ha.template to<utils::Location::Device>();
    const auto db = hb.template to<utils::Location::Device>();
    const auto ref = host::gustavson(ha, hb);
    const auto absref = host::abs_product(ha, hb);
    int bad = 0;
    for (int r = 0; r < repeat; ++r) {
        auto hc = graphs ? [&] {
            auto hs = ha;
            for (int i = 0; i < hs.nnz; ++i) hs.val[i] *= static_cast<T>(r + 1);
            const auto ds = hs.template to<utils::Location::Device>();
            auto out = proposal::proposal(ds, db).template to<utils::Location::Host>();
            for (int i = 0; i < out.nnz; ++i) out.val[i] /= static_cast<T>(r + 1);
            return out;
        }() : proposal::proposal(da, db).template to<utils::Location::Host>();
        const auto e = host::errors(hc, ref, absref);
        const double tol = sizeof(T) == 8 ? 1e-12 : 1e-5;
        std::printf("run %d: nnz %d (ref %d) structure %s scaled_err %.3e\n", r, hc.nnz, ref.nnz,
                    e.structure ? "identical" : "DIFFERENT", e.scaled);
        if (!e.structure || e.scaled > tol) bad = 1;
    }
    if (graphs) {
        int64_t captured = 0, replayed = 0;
        utils::handle_spg_error(spg_graph_stats(&captured, &replayed), "graph_stats");
        std::printf("graphs captured %lld replayed %lld\n", (long long)captured, (long long)replayed);
    }
    return bad;
}

// spg_multiply_noalloc into caller arrays: a capacity one short of nnz(C)
// must fail with SPG_ERR_OUT_OF_MEMORY and report the nnz needed; then
// `repeat` products at the exact capacity, each compared with the host.
template <typename T>
static int noalloc(const std::string& fa, const std::string& fb, int repeat) {
    const auto ha = utils::HostCSR<T>::load_from_filename(fa);
    const auto hb = utils::HostCSR<T>::load_from_filename(fb);
    const auto da = ha.template to<utils::Location::Device>();
    const auto db = hb.template to<utils::Location::Device>();
    const auto ref = host::gustavson(ha, hb);
    const auto absref = host::abs_product(ha, hb);
    const spg_dtype dt = sizeof(T) == 8 ? SPG_F64 : SPG_F32;
    const spg_csr A = da.raw(), B = db.raw();
    void *rpt = nullptr, *col = nullptr, *val = nullptr;
    utils::handle_spg_error(spg_malloc(&rpt, sizeof(int32_t) * (ha.m + 1)), "malloc");
    utils::handle_spg_error(spg_malloc(&col, sizeof(int32_t) * (ref.nnz + 1)), "malloc");
    utils::handle_spg_error(spg_malloc(&val, sizeof(T) * (ref.nnz + 1)), "malloc");
    int bad = 0;
    if (ref.nnz > 0) {
        spg_csr c{0, 0, 0, static_cast<int32_t*>(rpt), static_cast<int32_t*>(col), val};
        const spg_status st = spg_multiply_noalloc(&A, &B, &c, ref.nnz - 1, dt);
        std::printf("capacity %d: status %s, nnz %d\n", ref.nnz - 1, spg_status_str(st), c.nnz);
        if (st != SPG_ERR_OUT_OF_MEMORY || c.nnz != ref.nnz) bad = 1;
    }
    for (int r = 0; r < repeat; ++r) {
        spg_csr c{0, 0, 0, static_cast<int32_t*>(rpt), static_cast<int32_t*>(col), val};
        utils::handle_spg_error(spg_multiply_noalloc(&A, &B, &c, ref.nnz, dt), "multiply_noalloc");
        utils::HostCSR<T> hc(c.nnz, c.m, c.n);
        utils::handle_spg_error(spg_memcpy(hc.rpt, c.rpt, sizeof(int32_t) * (c.m + 1)), "memcpy");
        utils::handle_spg_error(spg_memcpy(hc.col, c.col, sizeof(int32_t) * c.nnz), "memcpy");
        utils::handle_spg_error(spg_memcpy(hc.val, c.val, sizeof(T) * c.nnz), "memcpy");
        const auto e = host::errors(hc, ref, absref);
        const double tol = sizeof(T) == 8 ? 1e-12 : 1e-5;
        std::printf("noalloc run %d: nnz %d (ref %d) structure %s scaled_err %.3e\n", r, hc.nnz, ref.nnz,
                    e.structure ? "identical" : "DIFFERENT", e.scaled);
        if (!e.structure || e.scaled > tol) bad = 1;
    }
    utils::handle_spg_error(spg_free(rpt), "free");
    utils::handle_spg_error(spg_free(col), "free");
    utils::handle_spg_error(spg_free(val), "free");
    return bad;
}

template <typename T>
static int serve(const std::string& fa, const std::string& fb, int n, long long budget, int max_batch) {
    const auto ha = utils::HostCSR<T>::load_from_filename(fa);
    const auto hb = utils::HostCSR<T>::load_from_filename(fb);
    const auto da = ha.template to<utils::Location::Device>();
    const auto db = hb.template to<utils::Location::Device>();
    const auto ref = host::gustavson(ha, hb);
    const auto absref = host::abs_product(ha, hb);
    const long long need = static_cast<long long>(ref.nnz) * (4 + sizeof(T));
    spg_server* srv = nullptr;
    utils::handle_spg_error(spg_server_create(static_cast<size_t>(budget), max_batch, &srv), "server_create");
    const spg_dtype dt = sizeof(T) == 8 ? SPG_F64 : SPG_F32;
    std::vector<int64_t> tickets(n);
    const spg_csr A = da.raw(), B = db.raw();
    for (int i = 0; i < n; ++i) utils::handle_spg_error(spg_server_submit(srv, &A, &B, dt, &tickets[i]), "submit");
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        spg_csr c{};
        double ms = 0;
        const spg_status st = spg_server_wait(srv, tickets[i], &c, &ms);
        if (need > budget) {
            std::printf("request %d: status %s (output %lld bytes over budget %lld)\n", i, spg_status_str(st), need,
                        budget);
            if (st != SPG_ERR_OUT_OF_MEMORY) bad = 1;
            continue;
        }
        if (st != SPG_OK) {
            std::printf("request %d failed: %s\n", i, spg_last_error());
            bad = 1;
            continue;
        }
        utils::DeviceCSR<T> dc;
        dc.nnz = c.nnz;
        dc.m = c.m;
        dc.n = c.n;
        dc.rpt = c.rpt;
        dc.col = c.col;
        dc.val = static_cast<T*>(c.val);
        dc.library_owned = true;
        const auto hc = dc.template to<utils::Location::Host>();
        const auto e = host::errors(hc, ref, absref);
        const double tol = sizeof(T) == 8 ? 1e-12 : 1e-5;
        std::printf("request %d: nnz %d structure %s scaled_err %.3e\n", i, hc.nnz,
                    e.structure ? "identical" : "DIFFERENT", e.scaled);
        if (!e.structure) {  // where it first differs
            int r = 0;
            while (r <= hc.m && hc.rpt[r] == ref.rpt[r]) ++r;
            int k = 0;
            while (k < std::min(hc.nnz, ref.nnz) && hc.col[k] == ref.col[k]) ++k;
            std::printf("  first row pointer differing: %d (%d vs %d); first column differing: %d\n", r,
                        r <= hc.m ? hc.rpt[r] : -1, r <= ref.m ? ref.rpt[r] : -1, k);
            std::printf("  columns from there:");
            for (int t = k; t < std::min(k + 8, hc.nnz); ++t) std::printf(" %d/%d", hc.col[t], ref.col[t]);
            int bad_cols = 0;
            for (int t = 0; t < hc.nnz; ++t) bad_cols += hc.col[t] != ref.col[t];
            std::printf("; %d of %d differ; C at rpt %p col %p val %p\n", bad_cols, hc.nnz, (void*)c.rpt,
                        (void*)c.col, c.val);
        }
        if (!e.structure || e.scaled > tol) bad = 1;
    }
    int64_t batches = 0, served = 0, max_seen = 0, deferred = 0, peak = 0;
    utils::handle_spg_error(spg_server_stats(srv, &batches, &served, &max_seen, &deferred, &peak), "stats");
    std::printf("served %lld in %lld batches (largest %lld, deferred %lld, peak output bytes %lld)\n",
                (long long)served, (long long)batches, (long long)max_seen, (long long)deferred, (long long)peak);
    if (peak > budget) bad = 1;
    utils::handle_spg_error(spg_server_destroy(srv), "server_destroy");
    return bad;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s A.csr B.csr [--f32] [--sym-table F] [--num-table F] [--repeat N]\n", argv[0]);
        return 2;
    }
    bool f32 = false, graphs = false, use_noalloc = false;
    int repeat = 1, serve_n = 0, max_batch = 1;
    long long budget = 0;
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--f32") {
            f32 = true;
        } else if ((a == "--sym-table" || a == "--num-table") && i + 1 < argc) {
            const std::string t = slurp(argv[++i]);
            utils::handle_spg_error(spg_set_tuning(a == "--sym-table" ? 0 : 1, t.c_str()), "set_tuning");
        } else if (a == "--graphs") {
            graphs = true;
        } else if (a == "--noalloc") {
            use_noalloc = true;
        } else if (a == "--repeat" && i + 1 < argc) {
            repeat = std::atoi(argv[++i]);
        } else if (a == "--serve" && i + 3 < argc) {
            serve_n = std::atoi(argv[++i]);
            budget = std::atoll(argv[++i]);
            max_batch = std::atoi(argv[++i]);
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        if (graphs) utils::handle_spg_error(spg_set_graphs(1), "set_graphs");
        if (serve_n > 0)
            return f32 ? serve<float>(argv[1], argv[2], serve_n, budget, max_batch)
                       : serve<double>(argv[1], argv[2], serve_n, budget, max_batch);
        if (use_noalloc)
            return f32 ? noalloc<float>(argv[1], argv[2], repeat) : noalloc<double>(argv[1], argv[2], repeat);
        return f32 ? run<float>(argv[1], argv[2], repeat, graphs) : run<double>(argv[1], argv[2], repeat, graphs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

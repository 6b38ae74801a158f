// Host Gustavson SpGEMM used by the apps' `validate` mode (the reference
// compares against cuSPARSE ALG2 there, src/apps/common.hpp:108-136).
// Row-parallel over the host's cores, dense accumulator + marker per
// thread, ascending columns, exact-zero sums kept as entries, products
// accumulated in double from 0.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "gpuspgemm.hpp"

namespace host {

template <typename Body>
void parallel_rows(std::int32_t m, Body body) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = m < 4096 ? 1u : hw;
    std::vector<std::thread> pool;
    const std::int64_t chunk = (static_cast<std::int64_t>(m) + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const std::int64_t lo = std::min<std::int64_t>(m, t * chunk), hi = std::min<std::int64_t>(m, lo + chunk);
        if (lo >= hi) continue;
        pool.emplace_back([=] { body(static_cast<std::int32_t>(lo), static_cast<std::int32_t>(hi)); });
    }
    for (auto& th : pool) th.join();
}

template <typename T>
utils::HostCSR<T> gustavson(const utils::HostCSR<T>& A, const utils::HostCSR<T>& B, bool absolute = false) {
    const std::int32_t m = A.m, n = B.n;
    std::vector<std::int64_t> cnt(static_cast<std::size_t>(m) + 1, 0);
    parallel_rows(m, [&](std::int32_t lo, std::int32_t hi) {
        std::vector<std::int32_t> mark(static_cast<std::size_t>(std::max(n, 1)), -1);
        for (std::int32_t i = lo; i < hi; ++i) {
            std::int64_t c = 0;
            for (std::int32_t p = A.rpt[i]; p < A.rpt[i + 1]; ++p)
                for (std::int32_t q = B.rpt[A.col[p]]; q < B.rpt[A.col[p] + 1]; ++q)
                    if (mark[B.col[q]] != i) {
                        mark[B.col[q]] = i;
                        ++c;
                    }
            cnt[static_cast<std::size_t>(i) + 1] = c;
        }
    });
    for (std::int32_t i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
    if (cnt[m] > INT32_MAX) throw std::runtime_error("host reference: nnz(C) does not fit int32");
    utils::HostCSR<T> C(static_cast<std::int32_t>(cnt[m]), m, n);
    for (std::int32_t i = 0; i <= m; ++i) C.rpt[i] = static_cast<std::int32_t>(cnt[i]);
    parallel_rows(m, [&](std::int32_t lo, std::int32_t hi) {
        std::vector<std::int32_t> mark(static_cast<std::size_t>(std::max(n, 1)), -1);
        std::vector<double> acc(static_cast<std::size_t>(std::max(n, 1)));
        for (std::int32_t i = lo; i < hi; ++i) {
            std::int32_t* cols = C.col + C.rpt[i];
            std::int32_t len = 0;
            for (std::int32_t p = A.rpt[i]; p < A.rpt[i + 1]; ++p) {
                const double a = A.val[p];
                for (std::int32_t q = B.rpt[A.col[p]]; q < B.rpt[A.col[p] + 1]; ++q) {
                    const std::int32_t j = B.col[q];
                    const double prod = a * static_cast<double>(B.val[q]);
                    if (mark[j] != i) {
                        mark[j] = i;
                        acc[j] = 0.0;
                        cols[len++] = j;
                    }
                    acc[j] += absolute ? std::fabs(prod) : prod;
                }
            }
            std::sort(cols, cols + len);
            for (std::int32_t t = 0; t < len; ++t) C.val[C.rpt[i] + t] = static_cast<T>(acc[cols[t]]);
        }
    });
    return C;
}

// (|A||B|)_ij on C's structure, for error scaling.
template <typename T>
utils::HostCSR<T> abs_product(const utils::HostCSR<T>& A, const utils::HostCSR<T>& B) {
    return gustavson(A, B, true);
}

struct ErrorStats {
    double elem = 0, scaled = 0;
    bool structure = false;
};

template <typename T>
ErrorStats errors(const utils::HostCSR<T>& C, const utils::HostCSR<T>& ref, const utils::HostCSR<T>& absref) {
    ErrorStats e;
    e.structure = C.m == ref.m && C.n == ref.n && C.nnz == ref.nnz &&
                  std::equal(C.rpt, C.rpt + C.m + 1, ref.rpt) && std::equal(C.col, C.col + C.nnz, ref.col);
    if (!e.structure) return e;
    for (std::int32_t t = 0; t < C.nnz; ++t) {
        const double d = std::fabs(static_cast<double>(C.val[t]) - static_cast<double>(ref.val[t]));
        if (ref.val[t] != 0) e.elem = std::max(e.elem, d / std::fabs(static_cast<double>(ref.val[t])));
        const double s = absref.val[t] > 0 ? static_cast<double>(absref.val[t]) : 1.0;
        e.scaled = std::max(e.scaled, d / s);
    }
    return e;
}

}  // namespace host

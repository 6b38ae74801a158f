// gpuspgemm.hpp -- C++ drop-in for the reference's host interface, built on
// the C ABI of gpuspgemm.h (no HIP header needed by users).
//
//   utils::CSR<T, Location>, HostCSR, DeviceCSR   csr.cuh:23-63, 254-258
//     load_from_filename / save_to_filename       csr.cuh:119-191
//     to<Location>()                              csr.cuh:193-203
//     operator==                                  csr.cuh:218-239
//   utils::get_nip(a, b)                          nip.cuh:52-78
//   utils::device_sync()                          runtime.cuh:98-101
//   proposal::proposal<T>(A, B) -> DeviceCSR<T>   proposal.cuh:85-124
//
// (paths relative to src/libutils/include/utils/ and
// src/libproposal/include/proposal/ of the reference).  Errors are thrown
// as std::runtime_error, as utils::handle_cuda_error does (errors.cuh:11-22).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gpuspgemm.h"

namespace utils {

enum class Location { Host, Device };

inline void handle_spg_error(spg_status status, const char* what) {
    if (status == SPG_OK) return;
    throw std::runtime_error(std::string("SpGEMM error (") + what + "): " + spg_status_str(status) + ": " +
                             spg_last_error());
}

template <Location L>
void* malloc(std::size_t bytes) {
    if constexpr (L == Location::Host) {
        void* p = std::malloc(bytes ? bytes : 1);
        if (!p) throw std::bad_alloc();
        return p;
    } else {
        void* p = nullptr;
        handle_spg_error(spg_malloc(&p, bytes), "malloc");
        return p;
    }
}

template <Location L>
void free(void* p) {
    if constexpr (L == Location::Host)
        std::free(p);
    else
        handle_spg_error(spg_free(p), "free");
}

inline void memcpy(void* dst, const void* src, std::size_t bytes) {
    handle_spg_error(spg_memcpy(dst, src, bytes), "memcpy");
}

inline void device_sync() { handle_spg_error(spg_device_sync(), "device_sync"); }

template <typename T>
constexpr spg_dtype dtype_of() {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double");
    return std::is_same_v<T, double> ? SPG_F64 : SPG_F32;
}

template <typename T, Location L>
struct CSR {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double");
    std::int32_t nnz{};
    std::int32_t m{};
    std::int32_t n{};
    std::int32_t* rpt{};
    std::int32_t* col{};
    T* val{};
    bool library_owned{};  // device arrays allocated by the multiply (released by spg_csr_free)

    CSR() = default;
    CSR(std::int32_t nnz_, std::int32_t m_, std::int32_t n_)
        : nnz{nnz_},
          m{m_},
          n{n_},
          rpt{static_cast<std::int32_t*>(malloc<L>((static_cast<std::size_t>(m_) + 1) * sizeof(std::int32_t)))},
          col{static_cast<std::int32_t*>(malloc<L>(static_cast<std::size_t>(nnz_) * sizeof(std::int32_t)))},
          val{static_cast<T*>(malloc<L>(static_cast<std::size_t>(nnz_) * sizeof(T)))} {}
    CSR(const CSR& o) : CSR(o.nnz, o.m, o.n) {
        memcpy(rpt, o.rpt, (static_cast<std::size_t>(m) + 1) * sizeof(std::int32_t));
        memcpy(col, o.col, static_cast<std::size_t>(nnz) * sizeof(std::int32_t));
        memcpy(val, o.val, static_cast<std::size_t>(nnz) * sizeof(T));
    }
    CSR(CSR&& o) noexcept
        : nnz{std::exchange(o.nnz, 0)},
          m{std::exchange(o.m, 0)},
          n{std::exchange(o.n, 0)},
          rpt{std::exchange(o.rpt, nullptr)},
          col{std::exchange(o.col, nullptr)},
          val{std::exchange(o.val, nullptr)},
          library_owned{std::exchange(o.library_owned, false)} {}
    CSR& operator=(CSR o) noexcept {
        swap(*this, o);
        return *this;
    }
    ~CSR() {
        try {
            release();
        } catch (...) {
        }
    }

    friend void swap(CSR& a, CSR& b) noexcept {
        using std::swap;
        swap(a.nnz, b.nnz);
        swap(a.m, b.m);
        swap(a.n, b.n);
        swap(a.rpt, b.rpt);
        swap(a.col, b.col);
        swap(a.val, b.val);
        swap(a.library_owned, b.library_owned);
    }

    void release() {
        if (library_owned) {
            spg_csr c{nnz, m, n, rpt, col, val};
            handle_spg_error(spg_csr_free(&c), "csr_free");
        } else {
            free<L>(rpt);
            free<L>(col);
            free<L>(val);
        }
        nnz = m = n = 0;
        rpt = col = nullptr;
        val = nullptr;
        library_owned = false;
    }

    // Little-endian: int32 nnz, m, n; int32 rpt[m+1]; int32 col[nnz];
    // float64 val[nnz] whatever T is (csr.cuh:137-154).
    static CSR load_from_filename(const std::string& filename) {
        static_assert(L == Location::Host, "files are read into host matrices");
        std::ifstream ifs(filename, std::ios::binary);
        if (!ifs) throw std::runtime_error("Failed to open file " + filename);
        const auto filesize = std::filesystem::file_size(filename);
        if (filesize < 3 * sizeof(std::int32_t)) throw std::runtime_error("File " + filename + " is too small");
        auto read = [&](void* data, std::uint64_t bytes) {
            ifs.read(static_cast<char*>(data), static_cast<std::streamsize>(bytes));
            if (ifs.fail()) throw std::runtime_error("Failed to read file " + filename);
        };
        std::int32_t hdr[3];
        read(hdr, sizeof(hdr));
        if (hdr[0] < 0 || hdr[1] < 0 || hdr[2] < 0) throw std::runtime_error("Failed to read file " + filename);
        CSR out(hdr[0], hdr[1], hdr[2]);
        read(out.rpt, (static_cast<std::uint64_t>(out.m) + 1) * sizeof(std::int32_t));
        read(out.col, static_cast<std::uint64_t>(out.nnz) * sizeof(std::int32_t));
        if constexpr (std::is_same_v<T, double>) {
            read(out.val, static_cast<std::uint64_t>(out.nnz) * sizeof(double));
        } else {
            std::vector<double> tmp(static_cast<std::size_t>(out.nnz));
            read(tmp.data(), tmp.size() * sizeof(double));
            std::copy(tmp.begin(), tmp.end(), out.val);
        }
        if (static_cast<std::uint64_t>(ifs.tellg()) < filesize)
            throw std::runtime_error("File " + filename + " is too large");
        return out;
    }

    void save_to_filename(const std::string& filename) const {
        static_assert(L == Location::Host, "files are written from host matrices");
        std::ofstream ofs(filename, std::ios::binary);
        if (!ofs) throw std::runtime_error("Failed to open file " + filename);
        auto write = [&](const void* data, std::uint64_t bytes) {
            ofs.write(static_cast<const char*>(data), static_cast<std::streamsize>(bytes));
            if (ofs.fail()) throw std::runtime_error("Failed to write to file");
        };
        const std::int32_t hdr[3] = {nnz, m, n};
        write(hdr, sizeof(hdr));
        write(rpt, (static_cast<std::uint64_t>(m) + 1) * sizeof(std::int32_t));
        write(col, static_cast<std::uint64_t>(nnz) * sizeof(std::int32_t));
        if constexpr (std::is_same_v<T, double>) {
            write(val, static_cast<std::uint64_t>(nnz) * sizeof(double));
        } else {
            std::vector<double> tmp(val, val + nnz);
            write(tmp.data(), tmp.size() * sizeof(double));
        }
    }

    template <Location To>
    CSR<T, To> to() const {
        static_assert(To != L, "Cannot convert to the same location");
        CSR<T, To> out(nnz, m, n);
        memcpy(out.rpt, rpt, (static_cast<std::size_t>(m) + 1) * sizeof(std::int32_t));
        memcpy(out.col, col, static_cast<std::size_t>(nnz) * sizeof(std::int32_t));
        memcpy(out.val, val, static_cast<std::size_t>(nnz) * sizeof(T));
        return out;
    }

    spg_csr raw() const { return spg_csr{nnz, m, n, rpt, col, val}; }
};

// Equal sizes, identical rpt and col, values equal or within
// atol 1e-8 + rtol 1e-5 * max(|l|, |r|) (csr.cuh:218-239).
template <typename T>
bool operator==(const CSR<T, Location::Host>& l, const CSR<T, Location::Host>& r) {
    if (&l == &r) return true;
    if (l.nnz != r.nnz || l.m != r.m || l.n != r.n) return false;
    if (!std::equal(l.rpt, l.rpt + l.m + 1, r.rpt)) return false;
    if (!std::equal(l.col, l.col + l.nnz, r.col)) return false;
    return std::equal(l.val, l.val + l.nnz, r.val, [](T a, T b) {
        constexpr T atol = T(1e-8);
        constexpr T rtol = T(1e-5);
        if (a == b) return true;
        return std::fabs(a - b) <= atol + rtol * std::max(std::fabs(a), std::fabs(b));
    });
}

template <typename T>
using HostCSR = CSR<T, Location::Host>;
template <typename T>
using DeviceCSR = CSR<T, Location::Device>;

template <typename T>
std::int64_t get_nip(const DeviceCSR<T>& a, const DeviceCSR<T>& b) {
    const spg_csr ra = a.raw(), rb = b.raw();
    std::int64_t nip = 0;
    handle_spg_error(spg_get_nip(&ra, &rb, &nip), "get_nip");
    return nip;
}

}  // namespace utils

namespace proposal {

// C = A . B; synchronous like the reference (C is complete on return).
template <typename T>
utils::DeviceCSR<T> proposal(const utils::DeviceCSR<T>& A, const utils::DeviceCSR<T>& B) {
    const spg_csr ra = A.raw(), rb = B.raw();
    spg_csr rc{};
    utils::handle_spg_error(spg_multiply(&ra, &rb, &rc, utils::dtype_of<T>()), "proposal");
    utils::DeviceCSR<T> C;
    C.nnz = rc.nnz;
    C.m = rc.m;
    C.n = rc.n;
    C.rpt = rc.rpt;
    C.col = rc.col;
    C.val = static_cast<T*>(rc.val);
    C.library_owned = true;
    return C;
}

}  // namespace proposal

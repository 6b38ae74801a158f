// Vendor-library comparator: C = A . B with rocSPARSE's generic SpGEMM,
// driven by the same app frame as `proposal` (the reference compares its
// method against cuSPARSE in src/apps/cusparse1.cu / cusparse2.cu through
// src/libutils/include/utils/cusparse.cuh; this is the MI355X analogue).
// Value type chosen at compile time with USE_DOUBLE_PRECISION.
//
// The timed region is the whole product as the library call sequence needs
// it (buffer-size query, buffer allocation, nnz stage, allocation of C,
// compute stage, synchronisation), like the `proposal` app's.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <chrono>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "gpuspgemm.hpp"

namespace {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

void check_rs(rocsparse_status s, const char* what) {
    if (s != rocsparse_status_success)
        throw std::runtime_error(std::string("rocSPARSE error ") + std::to_string(static_cast<int>(s)) + " in " + what);
}

template <typename T>
constexpr rocsparse_datatype datatype() {
    return sizeof(T) == 8 ? rocsparse_datatype_f64_r : rocsparse_datatype_f32_r;
}

struct Handle {
    rocsparse_handle h{};
    Handle() { check_rs(rocsparse_create_handle(&h), "create_handle"); }
    ~Handle() { (void)rocsparse_destroy_handle(h); }
};

template <typename T>
rocsparse_spmat_descr describe(const utils::DeviceCSR<T>& M) {
    rocsparse_spmat_descr d{};
    check_rs(rocsparse_create_csr_descr(&d, M.m, M.n, M.nnz, M.rpt, M.col, M.val, rocsparse_indextype_i32,
                                        rocsparse_indextype_i32, rocsparse_index_base_zero, datatype<T>()),
             "create_csr_descr");
    return d;
}

// C = A . B; C's arrays come from hipMalloc and are released by the CSR
// destructor (utils::free<Device> -> spg_free -> hipFree).
template <typename T>
utils::DeviceCSR<T> multiply(rocsparse_handle h, const utils::DeviceCSR<T>& A, const utils::DeviceCSR<T>& B) {
    const T alpha = 1, beta = 0;
    utils::DeviceCSR<T> C;
    C.m = A.m;
    C.n = B.n;
    C.rpt = static_cast<std::int32_t*>(utils::malloc<utils::Location::Device>(sizeof(std::int32_t) * (A.m + 1)));
    rocsparse_spmat_descr da = describe(A), db = describe(B), dc{}, dd{};
    check_rs(rocsparse_create_csr_descr(&dc, A.m, B.n, 0, C.rpt, nullptr, nullptr, rocsparse_indextype_i32,
                                        rocsparse_indextype_i32, rocsparse_index_base_zero, datatype<T>()),
             "create C");
    // D takes no part (beta = 0); it shares C's empty pattern
    check_rs(rocsparse_create_csr_descr(&dd, A.m, B.n, 0, C.rpt, nullptr, nullptr, rocsparse_indextype_i32,
                                        rocsparse_indextype_i32, rocsparse_index_base_zero, datatype<T>()),
             "create D");
    size_t bytes = 0;
    void* buf = nullptr;
    auto stage = [&](rocsparse_spgemm_stage st, const char* what) {
        check_rs(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, da, db, &beta, dd, dc,
                                  datatype<T>(), rocsparse_spgemm_alg_default, st, &bytes, buf),
                 what);
    };
    stage(rocsparse_spgemm_stage_buffer_size, "spgemm buffer size");
    check_hip(hipMalloc(&buf, bytes > 0 ? bytes : 1), "hipMalloc buffer");
    stage(rocsparse_spgemm_stage_nnz, "spgemm nnz");
    int64_t rows = 0, cols = 0, nnz = 0;
    check_rs(rocsparse_spmat_get_size(dc, &rows, &cols, &nnz), "spmat_get_size");
    C.nnz = static_cast<std::int32_t>(nnz);
    C.col = static_cast<std::int32_t*>(utils::malloc<utils::Location::Device>(sizeof(std::int32_t) * nnz));
    C.val = static_cast<T*>(utils::malloc<utils::Location::Device>(sizeof(T) * nnz));
    check_rs(rocsparse_csr_set_pointers(dc, C.rpt, C.col, C.val), "csr_set_pointers");
    stage(rocsparse_spgemm_stage_compute, "spgemm compute");
    check_hip(hipDeviceSynchronize(), "synchronize");
    check_hip(hipFree(buf), "hipFree buffer");
    for (auto d : {da, db, dc, dd}) (void)rocsparse_destroy_spmat_descr(d);
    return C;
}

}  // namespace

int main(int argc, char** argv) {
    using Clock = std::chrono::steady_clock;
#ifdef USE_DOUBLE_PRECISION
    using ValueType = double;
#else
    using ValueType = float;
#endif
    Handle handle;
    auto run = [&](const utils::DeviceCSR<ValueType>& d_a,
                   const utils::DeviceCSR<ValueType>& d_b) -> utils::HostCSR<ValueType> {
        return multiply(handle.h, d_a, d_b).template to<utils::Location::Host>();
    };
    auto measure = [&](const utils::DeviceCSR<ValueType>& d_a,
                       const utils::DeviceCSR<ValueType>& d_b) -> Clock::duration {
        const auto start = Clock::now();
        auto c = multiply(handle.h, d_a, d_b);
        utils::device_sync();
        const auto end = Clock::now();
        return end - start;
    };
    return app::run_app<ValueType>("rocSPARSE", run, measure, argc, argv);
}

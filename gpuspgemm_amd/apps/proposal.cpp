// The `proposal` app (src/apps/proposal.cu:6-44): one multiply per call on
// the GPU through the C ABI.  Value type chosen at compile time with
// USE_DOUBLE_PRECISION, as in the reference (cmake/options.cmake:2-3).
#include <chrono>

#include "common.hpp"
#include "gpuspgemm.hpp"

int main(int argc, char** argv) {
    using Clock = std::chrono::steady_clock;
#ifdef USE_DOUBLE_PRECISION
    using ValueType = double;
#else
    using ValueType = float;
#endif
    // multiply and bring C back to the host
    auto run = [](const utils::DeviceCSR<ValueType>& d_a,
                  const utils::DeviceCSR<ValueType>& d_b) -> utils::HostCSR<ValueType> {
        return proposal::proposal(d_a, d_b).template to<utils::Location::Host>();
    };
    // wall clock around the whole call + device synchronisation; C is freed
    // after the clock stops (src/apps/proposal.cu:28-35)
    auto measure = [](const utils::DeviceCSR<ValueType>& d_a,
                      const utils::DeviceCSR<ValueType>& d_b) -> Clock::duration {
        const auto start = Clock::now();
        auto c = proposal::proposal(d_a, d_b);
        utils::device_sync();
        const auto end = Clock::now();
        return end - start;
    };
    return app::run_app<ValueType>("Proposal", run, measure, argc, argv);
}

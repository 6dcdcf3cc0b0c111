// Single-qubit Kraus channels and the all-qubit noise model
// (reference include/lqcs/backend/noise.hpp:11-44).
#pragma once

#include <array>
#include <vector>

#include "../core/types.hpp"

namespace lqcs {

struct KrausChannel {
    std::vector<std::array<complex_t, 4>> ops;

    // Checks completeness sum_i K_i^dag K_i = I to 1e-10.
    explicit KrausChannel(std::vector<std::array<complex_t, 4>> kraus_ops);

    static KrausChannel depolarizing(double p);
    static KrausChannel amplitude_damping(double gamma);
    static KrausChannel phase_damping(double gamma);
    static KrausChannel bit_flip(double p);
    static KrausChannel phase_flip(double p);
};

// After every unitary instruction, each channel acts on each qubit the
// instruction touched, channel-major.
class NoiseModel {
public:
    NoiseModel() = default;
    NoiseModel& add_all_qubit_channel(KrausChannel channel) {
        channels_.push_back(std::move(channel));
        return *this;
    }
    const std::vector<KrausChannel>& channels() const { return channels_; }
    bool empty() const { return channels_.empty(); }

private:
    std::vector<KrausChannel> channels_;
};

}  // namespace lqcs

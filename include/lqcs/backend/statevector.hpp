// Statevector: 2^n complex128 amplitudes resident in HBM, initialised to
// |0...0> (reference include/lqcs/backend/statevector.hpp:12-37).
//
// data() / operator[] expose a host mirror that is synchronised on demand:
// reading downloads the device register when it is newer; writable access
// marks the mirror as newer so the next device operation uploads it.
#pragma once

#include <cstddef>
#include <string_view>
#include <vector>

#include "../core/types.hpp"

struct lqcs_dev_state;

namespace lqcs {

class Statevector {
public:
    static constexpr std::size_t kMaxQubits = 30;

    explicit Statevector(std::size_t num_qubits);
    // Register wider than kMaxQubits (up to the device limit), for callers
    // that opt in through EngineOptions::max_qubits.
    Statevector(std::size_t num_qubits, std::size_t max_qubits);
    ~Statevector();
    Statevector(const Statevector& other);
    Statevector& operator=(const Statevector& other);
    Statevector(Statevector&& other) noexcept;
    Statevector& operator=(Statevector&& other) noexcept;

    std::size_t num_qubits() const { return num_qubits_; }
    std::size_t size() const { return std::size_t{1} << num_qubits_; }

    complex_t& operator[](std::size_t i) { return data()[i]; }
    const complex_t& operator[](std::size_t i) const { return data()[i]; }
    complex_t* data();
    const complex_t* data() const;

    double norm() const;
    std::vector<double> probabilities() const;
    double expectation_value(std::string_view pauli) const;

    // Engine extension: reinitialise the register to |index>.
    void set_basis(std::size_t index);
    // Engine extension: max_i |a_i - other_i|, computed on the device.
    double max_abs_diff(const Statevector& other) const;

    // Engine access: the device register, made current first.
    lqcs_dev_state* device() const;
    // Declares that device code changed the register.
    void mark_device_newer();

private:
    enum class Fresh { Both, Host, Device };
    void make_host_current() const;

    std::size_t num_qubits_;
    lqcs_dev_state* dev_ = nullptr;
    mutable aligned_vector<complex_t> host_;
    mutable Fresh fresh_ = Fresh::Device;
    // Set once writable host access was handed out: the mirror may change at
    // any time after that, so every device use uploads it first, until device
    // code changes the register (mark_device_newer), which makes those views
    // stale.
    bool host_exported_ = false;
};

}  // namespace lqcs

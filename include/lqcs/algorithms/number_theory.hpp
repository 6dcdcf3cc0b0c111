// Classical number theory used by Shor (declarations as in the reference's
// include/lqcs/algorithms/number_theory.hpp:7-27; implemented by the
// reference's src/algorithms/number_theory.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace lqcs::algorithms::number_theory {

std::uint64_t gcd(std::uint64_t a, std::uint64_t b);
std::uint64_t mul_mod(std::uint64_t a, std::uint64_t b, std::uint64_t mod);  // no overflow
std::uint64_t pow_mod(std::uint64_t base, std::uint64_t exp, std::uint64_t mod);
bool is_prime(std::uint64_t n);  // trial division
// b when n = b^k for some k >= 2, else 0.
std::uint64_t perfect_power_base(std::uint64_t n);
// Denominators of the convergents of num/den up to max_denominator.
std::vector<std::uint64_t> continued_fraction_denominators(std::uint64_t num, std::uint64_t den,
                                                           std::uint64_t max_denominator);

}  // namespace lqcs::algorithms::number_theory

// Umbrella header of the algorithm library (the reference's
// include/lqcs/algorithms/algorithms.hpp:5-15 includes the same set).
#pragma once

#include "arithmetic.hpp"
#include "bernstein_vazirani.hpp"
#include "deutsch_jozsa.hpp"
#include "grover.hpp"
#include "number_theory.hpp"
#include "oracles.hpp"
#include "qft.hpp"
#include "qpe.hpp"
#include "shor.hpp"
#include "simon.hpp"
#include "vqe.hpp"

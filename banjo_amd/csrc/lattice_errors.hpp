// Error plumbing of the C ABI: C++ exceptions inside the library, integer
// codes plus a per-thread message at the boundary. The exception kinds are
// the reference's (std::runtime_error for no device or device failure,
// std::invalid_argument for bad input or an unsupported feature,
// std::overflow_error for ledger overflow, std::logic_error for misuse;
// src/fastlattice/FastLattice.hpp:400-406, CudaLatticeUnavailable.cpp:19-25).
#pragma once

#include "banjo_lattice.h"

#include <stdexcept>
#include <string>

namespace blx {

struct Runtime : std::runtime_error { using std::runtime_error::runtime_error; };
struct Invalid : std::invalid_argument { using std::invalid_argument::invalid_argument; };
struct Overflow : std::overflow_error { using std::overflow_error::overflow_error; };
struct Logic : std::logic_error { using std::logic_error::logic_error; };

std::string &lastError();

template <class F> int guard(F &&f) {
    try {
        f();
        lastError().clear();
        return BLX_OK;
    } catch (const std::invalid_argument &e) {
        lastError() = e.what();
        return BLX_ERR_INVALID;
    } catch (const std::overflow_error &e) {
        lastError() = e.what();
        return BLX_ERR_OVERFLOW;
    } catch (const std::logic_error &e) {
        lastError() = e.what();
        return BLX_ERR_LOGIC;
    } catch (const std::exception &e) {
        lastError() = e.what();
        return BLX_ERR_RUNTIME;
    } catch (...) {
        lastError() = "unknown failure";
        return BLX_ERR_RUNTIME;
    }
}

} // namespace blx

// guard.hpp — the C ABI's exception boundary (host code only).
#pragma once
#include "../../include/accspmm.h"

namespace accspmm {

// Runs the body of a C entry point: a C++ exception thrown inside (a host
// allocation that failed, a thread that could not be started) becomes
// ACCSPMM_ERR_OUT_OF_MEMORY, so none crosses the C ABI.
template <class F>
accspmm_status guard(F&& body) noexcept {
  try {
    return body();
  } catch (...) {
    return ACCSPMM_ERR_OUT_OF_MEMORY;
  }
}

}  // namespace accspmm

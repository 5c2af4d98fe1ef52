// dc_guard.h -- keeps C++ exceptions from crossing the extern "C" boundary.
//
// Every entry point of include/dyncolor.h whose body can allocate host memory
// (std::vector, std::string, new) runs inside guarded(): std::bad_alloc
// becomes DC_ERR_NOMEM, any other exception DC_ERR_INVALID, each with a
// message for dc_last_error().
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/dyncolor.h"

namespace dc {

void set_error(const std::string& msg);

template <class F>
int guarded(const char* where, F&& f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    try {
      set_error(std::string(where) + ": host allocation failed");
    } catch (...) {
    }
    return DC_ERR_NOMEM;
  } catch (const std::exception& e) {
    try {
      set_error(std::string(where) + ": " + e.what());
    } catch (...) {
    }
    return DC_ERR_INVALID;
  } catch (...) {
    try {
      set_error(std::string(where) + ": unknown exception");
    } catch (...) {
    }
    return DC_ERR_INVALID;
  }
}

}  // namespace dc

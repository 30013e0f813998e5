// vibestran_amd/csrc/vbh_error.hpp -- error reporting shared by the HIP and
// the host-only sources of libvbh: C-ABI entry points run their body through
// guarded(), which turns an Error (or a standard exception) into a status code
// and the thread-local message of vbh_last_error().
#pragma once
#include <exception>
#include <new>
#include <string>

#include "../../include/vbh.h"

namespace vbh {

struct Error {
  int code;
  std::string msg;
};
void set_error(const std::string& msg);
[[noreturn]] void throw_error(int code, const std::string& msg);
std::string fmt_g(double v);  // printf %g

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return VBH_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return VBH_ERR_HIP;
  } catch (const std::exception& e) {
    set_error(e.what());
    return VBH_ERR_ARG;
  }
}

}  // namespace vbh

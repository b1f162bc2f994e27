#include "status.hpp"

#include <cstdio>

namespace momms {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int bad_arg(const char* fn, int i, const char* what) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: argument %d (%s) is invalid", fn, i, what);
  g_last_error = buf;
  return -i;
}

int hip_status(const char* fn, hipError_t err) {
  if (err == hipSuccess) return 0;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: HIP error %d (%s)", fn, (int)err, hipGetErrorString(err));
  g_last_error = buf;
  return (int)err;
}

int check_operand(const char* fn, int arg_index, const void* p, int64_t h, int64_t w,
                  int64_t rs, int64_t cs) {
  if (h > 0 && w > 0 && p == nullptr) return bad_arg(fn, arg_index, "null pointer");
  if (rs < 1) return bad_arg(fn, arg_index + 1, "row stride < 1");
  if (cs < 1) return bad_arg(fn, arg_index + 2, "column stride < 1");
  return 0;
}

}  // namespace momms

extern "C" const char* momms_last_error_string(void) { return momms::g_last_error.c_str(); }

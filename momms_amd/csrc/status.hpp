// Status codes and the thread-local "last error" text behind momms_last_error_string().
//
// Convention (xerbla-style, like the Fortran dgemm_ the reference binds in
// src/util.rs:14-22): 0 = success, -i = argument i (1-based) is invalid,
// +hipError_t = a HIP runtime failure. No entry point aborts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace momms {

void set_error(const std::string& msg);
void clear_error();

// Records "name: argument i (what) is invalid" and returns -i.
int bad_arg(const char* fn, int i, const char* what);

// Records the HIP error text and returns +err (or 0 when err == hipSuccess).
int hip_status(const char* fn, hipError_t err);

// Checks the strides of an h x w operand: both positive, and at least one of
// them is 1 or the operand is a single row/column. Returns 0 or -arg_index.
int check_operand(const char* fn, int arg_index, const void* p, int64_t h, int64_t w,
                  int64_t rs, int64_t cs);

}  // namespace momms

#define MOMMS_HIP_TRY(fn, expr)                          \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return ::momms::hip_status(fn, _e); \
  } while (0)

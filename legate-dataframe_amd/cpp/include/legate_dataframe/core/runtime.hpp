/*
 * Per-process settings of the host API: the HIP stream every operation is enqueued on (the
 * reference runs each task on TaskContext::stream(), cpp/include/legate_dataframe/core/task_context.hpp:92-106)
 * and the exception mapping of C ABI status codes.
 */
#pragma once

#include <stdexcept>
#include <string>

#include <ldf_hip.h>

namespace legate::dataframe {

/* The stream used by every operation of this thread (hipStream_t; nullptr = the null stream). */
void set_stream(void* stream);
void* get_stream();

/* Raised for C ABI errors other than invalid arguments (those raise std::invalid_argument). */
struct HipError : std::runtime_error {
  int code;
  HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace detail {
/* Throws the exception the reference raises for the status code. */
void check_status(int rc);
void* device_alloc(size_t bytes);
void device_free(void* p);
void copy_to_device(void* dst, const void* src, size_t bytes);
void copy_to_host(void* dst, const void* src, size_t bytes);
}  // namespace detail

}  // namespace legate::dataframe

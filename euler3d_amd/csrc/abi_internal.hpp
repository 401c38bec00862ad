// Shared helpers of the C ABI implementation.
#pragma once
#include <string>

struct e3d_ctx;

namespace e3d {
void set_thread_error(const std::string& s);
// Records the message on the context (if any) and for the calling thread; returns code.
int set_error(e3d_ctx* ctx, int code, const std::string& msg);
void ctx_record_error(e3d_ctx* ctx, const std::string& msg);
const char* ctx_error_text(const e3d_ctx* ctx);
}  // namespace e3d

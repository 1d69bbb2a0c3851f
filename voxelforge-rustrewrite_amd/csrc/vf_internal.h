// Internal helpers shared by the library's translation units (not part of the C ABI).
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

struct vf_ctx;

namespace vf {

// Thread-local last error for calls without a context (and a copy for every context call).
std::string &thread_error();
void set_ctx_error(vf_ctx *ctx, const std::string &msg);

inline int fail(vf_ctx *ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    thread_error() = buf;
    if (ctx) set_ctx_error(ctx, buf);
    return code;
}

}  // namespace vf

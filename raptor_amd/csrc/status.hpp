// Error plumbing shared by the HIP and the host-only translation units:
// failures are thrown inside the library and become status codes at the C
// boundary, with the message kept per thread for ramd_last_error().
#pragma once

#include <new>

#include "raptor_amd.h"

namespace ramd {

struct Failure {
    int code;
};

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
[[noreturn]] void fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define RAMD_CHECK(cond, code, ...)                                                       \
    do {                                                                                  \
        if (!(cond)) ::ramd::fail(code, __VA_ARGS__);                                     \
    } while (0)

// Runs f() and converts a Failure (or an allocation failure) into a status code.
template <class F>
int guarded(F&& f)
{
    try {
        f();
        return RAMD_OK;
    } catch (const Failure& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return RAMD_E_ALLOC;
    } catch (const std::exception& e) {
        set_error("%s", e.what());
        return RAMD_E_ARG;
    }
}

}  // namespace ramd

// Host-only harness for p2g::guarded, the exception barrier around every
// int-returning C entry point: prints the status each kind of failure maps to.
#include <cstdio>
#include <stdexcept>
#include <system_error>

#include "p2g_context.hpp"

int main() {
    using namespace p2g;
    const int ok = guarded(nullptr, []() -> int { return 7; });
    const int oom = guarded(nullptr, []() -> int { throw std::bad_alloc(); });
    const int sys = guarded(nullptr, []() -> int {
        throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again));
    });
    const int other = guarded(nullptr, []() -> int { throw 3; });
    printf("%d %d %d %d\n", ok, oom, sys, other);
    return 0;
}

// common.hpp -- status codes, thread-local error text, small host helpers.
#pragma once
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cufhedb_amd.h"

namespace fhe {

void set_error(const char *fmt, ...);

#define FHE_CHECK_ARG(cond, ...)             \
    do {                                      \
        if (!(cond)) {                        \
            ::fhe::set_error(__VA_ARGS__);    \
            return FHE_ERR_ARG;               \
        }                                     \
    } while (0)

// Run `fn` over [0, n) on up to `threads` host threads (static blocking).
void parallel_for(int64_t n, int threads, void (*fn)(int64_t, int64_t, void *), void *arg);
int host_threads();

}  // namespace fhe

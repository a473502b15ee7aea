// Shared helpers of the C++ drop-in translation units: one verb_ctx per
// calling thread (the reference calls the operators concurrently from
// OpenMP workers) and the reference's fatal-error behaviour for a nonzero
// status (Logger::error + exit(EXIT_FAILURE), src/Logger.cpp:45-56).
#pragma once
#include <cstdlib>
#include <cstring>

#include "Logger.h"
#include "verb4d_hip.h"

namespace verb_dropin {

inline verb_ctx* ctx() {
    thread_local struct Holder {
        verb_ctx* c = nullptr;
        ~Holder() {
            if (c) verb_ctx_destroy(c);
        }
    } h;
    if (!h.c) {
        int rc = verb_ctx_create(&h.c);
        if (rc != 0) {
            Logger::error << "verb4d: no HIP device for the GPU operators (status " << rc << ")" << std::endl;
            exit(EXIT_FAILURE);
        }
    }
    return h.c;
}

inline void check(int rc, const char* what) {
    if (rc == 0) return;
    Logger::error << what << ": " << verb_ctx_last_error(ctx()) << " (status " << rc << ")" << std::endl;
    exit(EXIT_FAILURE);
}

}  // namespace verb_dropin

// Phase timer of the compiler passes, printed to stderr when SYMLIB_TIMING is set (the
// reference times its phases through Timer, src/Timer.cpp:71-98).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace Sym {

struct PhaseTimer {
    bool on = std::getenv("SYMLIB_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void operator()(const char* what) {
        if (!on) return;
        auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[symlib] %-28s %8.3f s\n", what, std::chrono::duration<double>(now - t).count());
        t = now;
    }
};

}  // namespace Sym

// timer.h -- wall-clock phase timer printing "<label>: <ms> ms" lines in the
// format of the reference's Timer (timer.h:13-75 there), on std::chrono.
#pragma once

#include <chrono>
#include <cstdio>

namespace cli {

enum PrintColor { NONE, GREEN, CYAN };

struct Timer {
    bool active = false;
    std::chrono::steady_clock::time_point t0, t1;
    explicit Timer(bool on) : active(on) {}
    void start() {
        if (active) t0 = std::chrono::steady_clock::now();
    }
    void printAndStart(const char* s) {
        if (active) {
            std::printf("%s", s);
            start();
        }
    }
    double elapsed_ms() const { return std::chrono::duration<double, std::milli>(t1 - t0).count(); }
    void stopAndPrint(const char* s = nullptr, PrintColor color = NONE) {
        if (!active) return;
        t1 = std::chrono::steady_clock::now();
        print(elapsed_ms(), s, color);
    }
    static void print(double ms, const char* s, PrintColor color = NONE) {
        if (color == GREEN) std::printf("\033[0;32m");
        if (color == CYAN) std::printf("\033[0;36m");
        if (s) std::printf("%s: %f ms\n", s, ms);
        else std::printf("%f ms\n", ms);
        if (color != NONE) std::printf("\033[0m");
    }
};

}  // namespace cli

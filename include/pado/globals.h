// pado/globals.h -- types and constants of the PADO C++ API, as the reference
// defines them (includes/globals.h:26-43), for the GPU drop-in.
#ifndef PADO_GPU_DROPIN_GLOBALS_H_
#define PADO_GPU_DROPIN_GLOBALS_H_

#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include <string>
#include <vector>

namespace PADO {

typedef uint32_t idi;          // vertex id
typedef uint8_t weighti;       // hop distance / edge weight
typedef int16_t weightiLarge;  // weighted distance (WPARA)
typedef uint8_t smalli;
typedef uint32_t inti;

const uint8_t WEIGHTI_MAX = UCHAR_MAX;
const int16_t WEIGHTILARGE_MAX = SHRT_MAX;
const uint8_t SMALLI_MAX = UCHAR_MAX;

// Printed by the reference's construct() only; the GPU build ignores it.
static inti NUM_THREADS = 4;

// Wall-clock timer with the reference's interface (includes/globals.h:84-121).
class WallTimer {
    double start_ = 0.0;
    std::string item_;
    static double now()
    {
        timeval t;
        gettimeofday(&t, nullptr);
        return t.tv_sec + t.tv_usec * 0.000001;
    }

public:
    WallTimer() : start_(now()) {}
    explicit WallTimer(const char *n) : start_(now()), item_(n) {}
    double get_runtime() { return now() - start_; }
    void print_runtime()
    {
        printf("%s: %f\n", item_.c_str(), get_runtime());
        fflush(stdout);
    }
    static double get_time_mark() { return now(); }
};

}  // namespace PADO

#endif  // PADO_GPU_DROPIN_GLOBALS_H_

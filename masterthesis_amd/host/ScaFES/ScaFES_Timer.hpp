// ScaFES_Timer.hpp -- host wall clock (role of ScaFES::Timer,
// src/ScaFES_Timer.hpp:171-214).  Device phases are timed after a stream
// synchronisation, so elapsed() covers the GPU work of the phase.
#ifndef SCAFES_HOST_TIMER_HPP_
#define SCAFES_HOST_TIMER_HPP_

#include <chrono>

namespace ScaFES {

class Timer {
public:
    Timer() : mStart(clock::now()) {}
    void restart() { mStart = clock::now(); }
    double elapsed() const { return std::chrono::duration<double>(clock::now() - mStart).count(); }

private:
    typedef std::chrono::steady_clock clock;
    clock::time_point mStart;
};

}  // namespace ScaFES
#endif

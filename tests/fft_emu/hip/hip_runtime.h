// Host stand-in for the few HIP names the FFT core (fft_core.hpp) uses, so
// that tests/fft_harness.cpp can run the device FFT stages on the CPU: one
// host thread per GPU thread, the workgroup barrier a std::barrier.  Test
// code only.
#pragma once
#include <barrier>

#define __device__
#define __host__
#define __forceinline__ inline

struct float2 {
  float x, y;
};
struct double2 {
  double x, y;
};
struct EmuDim3 {
  unsigned x, y, z;
};
extern thread_local EmuDim3 threadIdx;
void __syncthreads();

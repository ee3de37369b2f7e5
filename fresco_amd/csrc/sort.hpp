// Stable radix sort of uint32 (key, value) pairs on the device, built from
// kernel launches only (no memory fills), so it can be captured in a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace fresco_hip {

// Records carried along by the sort: element i of the output gets
// rec_in[v] (rec_words 32-bit words, a multiple of 4) and cell_in[v], v the
// output value (the original index).  Moved in the last digit pass only.
struct SortPayload {
  const void* rec_in = nullptr;
  void* rec_out = nullptr;
  int rec_words = 4;
  const uint32_t* cell_in = nullptr;
  uint32_t* cell_out = nullptr;
};

class BinSort {
 public:
  // n pairs; keys are below 2^key_bits
  void init(int64_t n, int key_bits);
  // keys / vals -> keys_out / vals_out (stable; inputs are not modified);
  // with a payload, its records are placed in the output order too
  void sort(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out, hipStream_t s,
            const SortPayload* payload = nullptr);

 private:
  int64_t n_ = 0;
  int passes_ = 1;
  int dbits_ = 1;  // digit width, at most 11 bits
  int nblk_ = 1;
  int nseg_ = 1;
  DeviceBuffer<uint32_t> cnt_, tops_, ktmp_, vtmp_;
};

}  // namespace fresco_hip
